"""The C ABI library and the CLI, without a GPU: the library loads, exports
every function include/nbx.h declares, reports errors through its status
codes, and the CLI keeps the reference's argument handling (Main.cpp:65-82)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nbx.h")
CLI = os.path.join(ROOT, "n-body-simulation_amd", "build", "nbx_sim")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree(nbx):
    assert declared_functions() == sorted(nbx.EXPORTS)


def test_library_exports_every_declared_symbol(nbx):
    L = nbx.load_library()
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert L.nbx_abi_version() == 1


def test_default_params_are_the_reference_defaults(nbx):
    L = nbx.load_library()
    p = nbx.Params()
    L.nbx_default_params(p)
    assert (p.dt, p.theta, p.epsilon, p.collisions) == (pytest.approx(0.05), 1.0, 1.0, 1)


def test_argument_errors_without_gpu(nbx):
    import ctypes as C
    L = nbx.load_library()
    assert L.nbx_create(0, C.byref(C.c_void_p())) == 1  # NBX_ERR_ARG
    assert b"capacity" in L.nbx_last_error()
    assert L.nbx_create(1, None) == 1
    assert L.nbx_destroy(None) == 0
    assert L.nbx_step(None, 1) == 1
    assert L.nbx_step_async(None, 1) == 1
    assert L.nbx_create_ex(16, 0x2, C.byref(C.c_void_p())) == 1  # unknown flag
    assert b"flag" in L.nbx_last_error()


@pytest.mark.parametrize("args", [[], ["10"], ["10", "5"], ["1", "2", "3", "4", "5"]])
def test_cli_usage(args):
    r = subprocess.run([CLI] + args, capture_output=True, text=True)
    assert r.returncode == 1
    assert "Usage:" in r.stderr and "<n_bodies> <total_steps> <output_every>" in r.stderr


def test_cli_output_every_zero():
    r = subprocess.run([CLI, "10", "5", "0"], capture_output=True, text=True)
    assert r.returncode == 1 and "output_every must be >= 1" in r.stderr
