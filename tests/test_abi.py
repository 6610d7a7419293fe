"""The C-ABI library loads and exports what include/bemhip.h declares (CPU)."""

import ctypes
import math
import os
import sys

import numpy as np

import bemhip


def test_library_exports_every_declared_symbol():
    names = bemhip.declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(bemhip.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_kernel_quadrature_constants_match_closed_forms(oracle_lib):
    q = bemhip.quadrature()
    o = oracle_lib.quadrature()
    for k in q:
        assert np.max(np.abs(q[k] - o[k])) < 3e-16, k
    s = math.sqrt(15.0)
    assert abs(q["tri_x"][1] - (9 + 2 * s) / 21) < 2e-16
    assert abs(q["tri_w"].sum() - 0.5) < 1e-15 and abs(q["gauss_w"].sum() - 1.0) < 1e-15


def test_no_silent_fallback_without_device():
    """On a host without a gfx950 device the context refuses to start."""
    import torch
    if torch.cuda.is_available():
        return
    try:
        bemhip.Context(0)
    except bemhip.BemHipError as e:
        assert e.code == bemhip.ERR_NO_DEVICE
    else:
        raise AssertionError("context created without a device")


def test_missing_library_fails_loudly(tmp_path):
    try:
        bemhip._lib_saved = bemhip._lib
        bemhip._lib = None
        bemhip.load(str(tmp_path / "nope.so"))
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("load of a missing library did not raise")
    finally:
        bemhip._lib = bemhip._lib_saved


def test_cpp_surface_compiles_and_links(tmp_path):
    """include/bemhip.hpp (the ColocSim-shaped C++ surface) builds against the library."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "use.cpp"
    src.write_text('#include "bemhip.hpp"\n'
                   "int main() {\n"
                   "  bemhip::Context c(0);\n"
                   "  std::vector<bemhip::vec3d> v{{0,0,0},{1,0,0},{0,1,0},{0,0,1}};\n"
                   "  std::vector<bemhip::tri3> t{{0,2,1},{0,1,3},{0,3,2},{1,2,3}};\n"
                   "  bemhip::ColocSim s(c, v, t, 1.0, 1.0, 0.0, 1.0);\n"
                   "  s.set_bcgstab(false); s.set_phi(1.0); s.evolve_system_RK4(1e-3, true);\n"
                   "  return s.get_psi().size() == 4 ? 0 : 1;\n}\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(root, "include"), str(src),
                        "-o", str(tmp_path / "use"), "-L", os.path.dirname(bemhip.LIB_PATH), "-lbemhip"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so,
    before any device work (runs on the CPU)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
