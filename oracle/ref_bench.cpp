// ref_bench.cpp -- CPU baseline driver over the reference sources
// (TEST/BENCH INFRASTRUCTURE). Built by oracle/Makefile into
// oracle/_ref/nbx_ref_bench with the reference's shipped optimisation flags
// and OpenMP. It runs the reference Simulation (Simulation.cpp:11-16) with
// the tree's theta set from the command line -- the shipped ./sim hard-codes
// theta = 1 (Simulation.cpp:7-9) -- and prints a line in the reference's
// "BH:" format (Main.cpp:210-216).
//
// usage: nbx_ref_bench <n_bodies> <steps> <theta> [warmup_steps]

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "Simulation.h"

int main(int argc, char **argv) {
  if (argc < 4 || argc > 5) {
    std::fprintf(stderr,
                 "usage: %s <n_bodies> <steps> <theta> [warmup_steps]\n",
                 argv[0]);
    return 1;
  }
  const std::size_t n = std::stoul(argv[1]);
  const unsigned steps = static_cast<unsigned>(std::stoul(argv[2]));
  const float theta = std::stof(argv[3]);
  const unsigned warmup = argc == 5 ? std::stoul(argv[4]) : 0;

  const auto setup0 = std::chrono::steady_clock::now();
  Simulation sim(n);
  sim.quadtree = QuadTree(theta, 1.0f, 16);
  const double setupMs = std::chrono::duration<double, std::milli>(
                             std::chrono::steady_clock::now() - setup0)
                             .count();
  for (unsigned k = 0; k < warmup; ++k) {
    sim.step();
  }
  double computeMs = 0.0;
  for (unsigned k = 0; k < steps; ++k) {
    const auto t0 = std::chrono::steady_clock::now();
    sim.step();
    computeMs += std::chrono::duration<double, std::milli>(
                     std::chrono::steady_clock::now() - t0)
                     .count();
    std::fprintf(stderr, "step %u/%u %.3f ms\n", k + 1, steps,
                 std::chrono::duration<double, std::milli>(
                     std::chrono::steady_clock::now() - t0)
                     .count());
  }
#ifdef _OPENMP
  const int threads = omp_get_max_threads();
#else
  const int threads = 1;
#endif
  std::printf("BH: n=%zu threads=%d steps=%u theta=%g warmup=%u setup_time=%.3f "
              "ms compute_time=%.3f ms avg_compute_step=%.4f ms\n",
              n, threads, steps, theta, warmup, setupMs, computeMs,
              steps ? computeMs / steps : 0.0);
  return 0;
}
