// nbx_cli.cpp -- `nbx_sim <n_bodies> <total_steps> <output_every> [output_path]`
//
// Drop-in for the reference's ./sim and ./sim_cuda (src/Main.cpp:65-219,
// src/cuda/MainCuda.cu:22-187): same arguments and errors, same NBS2 file
// (28-byte header with a live completed_frames counter, mass and radius once,
// then [x, y, vx, vy] frames in id order; frame 0 is the initial state), and
// a stderr summary line starting with "BH: " that cluster/times.sh parses.
//
// The physics parameters the reference fixes at compile time can be set
// from the environment: NBX_THETA, NBX_EPSILON, NBX_DT, NBX_COLLISIONS (0/1),
// NBX_LAW (reference|accurate|quadrupole), NBX_SUM (reference|compensated), NBX_COLLIDE
// (sequential|jacobi: the reference CPU's or GPU's resolution). NBX_SCENARIO
// (disc|two-galaxies|cold-square|merger) and NBX_SEED pick the initial
// condition (default: the reference's uniformDisc, seed 0).

#include "nbx.h"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

void die(const char *what) {
  std::cerr << what << ": " << nbx_last_error() << "\n";
  std::exit(1);
}

float envFloat(const char *name, float dflt) {
  const char *v = std::getenv(name);
  return v ? std::stof(v) : dflt;
}

} // namespace

int main(int argc, char **argv) {
  if (argc != 4 && argc != 5) {
    std::cerr << "Usage: " << argv[0]
              << " <n_bodies> <total_steps> <output_every> [output_path]\n"
              << "  (no output_path: run without writing any output)\n";
    return 1;
  }
  const std::size_t nBodies = static_cast<std::size_t>(std::stoul(argv[1]));
  const std::uint32_t totalSteps = static_cast<std::uint32_t>(std::stoul(argv[2]));
  const std::uint32_t outputEvery = static_cast<std::uint32_t>(std::stoul(argv[3]));
  const bool writeOutput = argc == 5;
  const std::string outputPath = writeOutput ? argv[4] : "";
  if (outputEvery == 0) {
    std::cerr << "output_every must be >= 1\n";
    return 1;
  }
  std::ofstream out;
  if (writeOutput) {
    out.open(outputPath, std::ios::binary);
    if (!out) {
      std::cerr << "Failed to open output file: " << outputPath << "\n";
      return 1;
    }
  }

  nbx_sim *sim = nullptr;
  if (nbx_create(nBodies > 0 ? nBodies : 1, &sim) != NBX_OK) {
    die("nbx_create");
  }
  nbx_params p;
  nbx_default_params(&p);
  p.theta = envFloat("NBX_THETA", p.theta);
  p.epsilon = envFloat("NBX_EPSILON", p.epsilon);
  p.dt = envFloat("NBX_DT", p.dt);
  if (const char *v = std::getenv("NBX_COLLISIONS")) {
    p.collisions = std::atoi(v) != 0;
  }
  if (const char *v = std::getenv("NBX_LAW")) {
    p.law = std::strcmp(v, "accurate") == 0     ? NBX_LAW_ACCURATE
            : std::strcmp(v, "quadrupole") == 0 ? NBX_LAW_ACCURATE_QUAD
                                                : NBX_LAW_REFERENCE;
  }
  if (const char *v = std::getenv("NBX_COLLIDE")) {
    p.collide_mode =
        std::strcmp(v, "jacobi") == 0 ? NBX_COLLIDE_JACOBI : NBX_COLLIDE_SEQUENTIAL;
  }
  if (const char *v = std::getenv("NBX_SUM")) {
    p.summation =
        std::strcmp(v, "reference") == 0 ? NBX_SUM_REFERENCE : NBX_SUM_COMPENSATED;
  }
  if (nbx_set_params(sim, &p) != NBX_OK) {
    die("nbx_set_params");
  }

  // uniformDisc, then into id order before the first step (Main.cpp:117-122).
  int scenario = NBX_SCENARIO_DISC;
  if (const char *v = std::getenv("NBX_SCENARIO")) {
    if (std::strcmp(v, "two-galaxies") == 0) {
      scenario = NBX_SCENARIO_TWO_GALAXIES;
    } else if (std::strcmp(v, "cold-square") == 0) {
      scenario = NBX_SCENARIO_COLD_SQUARE;
    } else if (std::strcmp(v, "merger") == 0) {
      scenario = NBX_SCENARIO_MERGER;
    } else if (std::strcmp(v, "disc") != 0) {
      std::fprintf(stderr, "unknown NBX_SCENARIO '%s'\n", v);
      return 1;
    }
  }
  const char *seedText = std::getenv("NBX_SEED");
  const std::uint32_t seed =
      seedText ? static_cast<std::uint32_t>(std::strtoul(seedText, nullptr, 10)) : 0u;
  if (nbx_init_scenario(sim, scenario, nBodies, seed) != NBX_OK) {
    die("nbx_init_scenario");
  }
  {
    std::vector<float> px(nBodies), py(nBodies), vx(nBodies), vy(nBodies),
        m(nBodies), r(nBodies);
    std::vector<std::uint32_t> id(nBodies);
    for (std::size_t i = 0; i < nBodies; ++i) {
      id[i] = static_cast<std::uint32_t>(i);
    }
    if (nbx_download_by_id(sim, px.data(), py.data(), vx.data(), vy.data(),
                           m.data(), r.data()) != NBX_OK ||
        nbx_upload(sim, nBodies, px.data(), py.data(), vx.data(), vy.data(),
                   nullptr, nullptr, m.data(), r.data(), id.data()) != NBX_OK) {
      die("initial id ordering");
    }
  }

  const std::uint32_t n32 = static_cast<std::uint32_t>(nBodies);
  const std::uint32_t numFrames = 1 + totalSteps / outputEvery;
  std::uint32_t completedFrames = 0;
  std::streampos completedOffset;
  std::vector<float> px(nBodies), py(nBodies), vx(nBodies), vy(nBodies),
      m(nBodies), r(nBodies);
  std::vector<float> frame(writeOutput ? nBodies * 4 : 0);
  if (writeOutput) {
    const char magic[4] = {'N', 'B', 'S', '2'};
    out.write(magic, 4);
    out.write(reinterpret_cast<const char *>(&n32), 4);
    out.write(reinterpret_cast<const char *>(&totalSteps), 4);
    out.write(reinterpret_cast<const char *>(&outputEvery), 4);
    out.write(reinterpret_cast<const char *>(&numFrames), 4);
    completedOffset = out.tellp();
    out.write(reinterpret_cast<const char *>(&completedFrames), 4);
    out.write(reinterpret_cast<const char *>(&p.dt), 4);
    if (nbx_download_by_id(sim, nullptr, nullptr, nullptr, nullptr, m.data(),
                           r.data()) != NBX_OK) {
      die("download");
    }
    out.write(reinterpret_cast<const char *>(m.data()),
              static_cast<std::streamsize>(nBodies * 4));
    out.write(reinterpret_cast<const char *>(r.data()),
              static_cast<std::streamsize>(nBodies * 4));
  }
  auto writeFrame = [&]() {
    if (nbx_download_by_id(sim, px.data(), py.data(), vx.data(), vy.data(),
                           nullptr, nullptr) != NBX_OK) {
      die("download");
    }
    for (std::size_t i = 0; i < nBodies; ++i) {
      frame[4 * i + 0] = px[i];
      frame[4 * i + 1] = py[i];
      frame[4 * i + 2] = vx[i];
      frame[4 * i + 3] = vy[i];
    }
    out.write(reinterpret_cast<const char *>(frame.data()),
              static_cast<std::streamsize>(frame.size() * 4));
    ++completedFrames;
    const std::streampos saved = out.tellp();
    out.seekp(completedOffset);
    out.write(reinterpret_cast<const char *>(&completedFrames), 4);
    out.seekp(saved);
  };
  if (writeOutput) {
    writeFrame();
  }

  // Compute time covers the queued steps up to each frame boundary (one
  // device sync per nbx_step call); output time is the rest of the loop.
  double computeMs = 0.0;
  const auto loopStart = std::chrono::steady_clock::now();
  std::uint32_t done = 0;
  while (done < totalSteps) {
    std::uint32_t chunk = totalSteps - done;
    if (writeOutput) {
      chunk = outputEvery - (done % outputEvery);
      if (chunk > totalSteps - done) {
        chunk = totalSteps - done;
      }
    }
    const auto t0 = std::chrono::steady_clock::now();
    if (nbx_step(sim, chunk) != NBX_OK) {
      die("nbx_step");
    }
    computeMs += std::chrono::duration<double, std::milli>(
                     std::chrono::steady_clock::now() - t0)
                     .count();
    done += chunk;
    if (writeOutput && done % outputEvery == 0) {
      writeFrame();
    }
  }
  const double loopMs = std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - loopStart)
                            .count();
  if (writeOutput) {
    out.close();
    if (!out) {
      std::cerr << "Error while writing output file: " << outputPath << "\n";
      return 1;
    }
  }
  nbx_destroy(sim);
  std::cerr << "Done. n=" << nBodies << " steps=" << totalSteps
            << " output_every=" << outputEvery
            << " frames_written=" << completedFrames << " (expected "
            << (writeOutput ? numFrames : 0) << ") -> "
            << (writeOutput ? outputPath : "(no output)") << "\n";
  std::cerr << "BH: n=" << nBodies << " threads=1 steps=" << totalSteps
            << " total_time=" << loopMs << " ms"
            << " avg_step=" << (totalSteps ? loopMs / totalSteps : 0.0)
            << " ms compute_time=" << computeMs << " ms"
            << " avg_compute_step="
            << (totalSteps ? computeMs / totalSteps : 0.0) << " ms"
            << " output_time=" << (loopMs - computeMs) << " ms"
            << " device=gfx950 theta=" << p.theta << "\n";
  return 0;
}
