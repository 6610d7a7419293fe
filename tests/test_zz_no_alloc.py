"""The no-allocation exterior potential (bemhip_exterior_potential_noalloc),
called through the C-ABI on caller-owned device buffers, against the
allocating entry point bemhip_exterior_potential on the same inputs: two calls
in a row, the second smaller, then the error paths (a size beyond the fixed
scratch array, a triangle index out of range found on the device).
Tolerance 1e-13 relative: the same kernels, only the triangle chunking of the
partial sums may differ.  The test's device buffers come from the HIP runtime
the library itself links (hipMalloc through ctypes), not from a second
runtime in the process."""

import ctypes as C
import os
import sys

import numpy as np
import pytest

import bemhip
import meshgen

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libbemhip links"""

    def __init__(self):
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        self.L = C.CDLL(path)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p.value)
        assert self.L.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value

    def get(self, ptr, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _case(nu, P, seed):
    v, t = meshgen.icosphere(nu, 1.0, (0.1, -0.2, 0.05))
    rng = np.random.default_rng(seed)
    phi = rng.standard_normal(len(v))
    psi = rng.standard_normal(len(v))
    d = rng.standard_normal((P, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = d * rng.uniform(0.2, 3.0, size=(P, 1))
    return v, t, phi, psi, pts


def test_exterior_noalloc_matches_allocating_entry():
    ctx = bemhip.Context(0)
    hip = _Hip()
    try:
        for nu, P, seed in ((24, 5000, 1), (8, 700, 2)):
            say(f"no-alloc exterior potential nu={nu} P={P}")
            v, t, phi, psi, pts = _case(nu, P, seed)
            ref = ctx.exterior_potential(v, t, phi, psi, pts)
            dv, dt = hip.put(v), hip.put(t.astype(np.uint32))
            dphi, dpsi, dpts = hip.put(phi), hip.put(psi), hip.put(pts)
            dout = hip.put(np.full(P, np.nan))
            ctx.exterior_potential_noalloc(len(v), dv, len(t), dt, dphi, dpsi, P, dpts, dout)
            got = hip.get(dout, P, np.float64)
            assert np.all(np.isfinite(got))
            assert rel(got, ref) < 1e-13
        # a size beyond the scratch array is refused before any device work
        with pytest.raises(bemhip.BemHipError) as e:
            ctx.exterior_potential_noalloc(len(v), dv, 100000, dt, dphi, dpsi, P, dpts, dout)
        assert e.value.code == bemhip.ERR_ARG
        # an index out of range is found by the device check (no kernel reads it)
        tb = t.astype(np.uint32).copy()
        tb[3, 1] = len(v) + 7
        dbad = hip.put(tb)
        with pytest.raises(bemhip.BemHipError) as e:
            ctx.exterior_potential_noalloc(len(v), dv, len(t), dbad, dphi, dpsi, P, dpts, dout)
        assert e.value.code == bemhip.ERR_BAD_MESH
        # and the entry point still works after the refusals
        ctx.exterior_potential_noalloc(len(v), dv, len(t), dt, dphi, dpsi, P, dpts, dout)
        assert rel(hip.get(dout, P, np.float64), ref) < 1e-13
    finally:
        hip.free()
        ctx.close()
