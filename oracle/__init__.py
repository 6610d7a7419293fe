"""ctypes front end of the FP64 CPU oracle (bem_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker used by tests/, the smoke entry point
and the benchmark's CPU-baseline leg.  The product path never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [
        ("p_inf", C.c_double), ("epsilon", C.c_double), ("sigma", C.c_double), ("gamma", C.c_double),
        ("V_0", C.c_double),
        ("field_amp", C.c_double), ("field_freq", C.c_double), ("field_phase", C.c_double),
        ("dp_balance", C.c_double), ("min_dt", C.c_double), ("wall_eps", C.c_double),
        ("mirror", C.c_int), ("wall_guards", C.c_int), ("frozen_stage_time", C.c_int),
        ("solver", C.c_int), ("nthreads", C.c_int), ("n_pin", C.c_int64),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(os.path.join(_HERE, "bem_oracle.c")):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = bind(_LIB)
    return _lib


def bind(path):
    """Loads an oracle build at `path` with its argument types (the mutation
    tests bind builds with planted mistakes this way)."""
    L = C.CDLL(path)
    L.bo_quadrature.argtypes = [f64p] * 5
    L.bo_panel.argtypes = [f64p, f64p, f64p, f64p, C.c_int, C.c_int, f64p, f64p]
    L.bo_assemble.argtypes = [C.c_int64, f64p, C.c_int64, u32p, C.c_int, f64p, f64p, C.c_int]
    L.bo_gemv.argtypes = [C.c_int64, f64p, f64p, f64p, C.c_int]
    L.bo_lu_factor.argtypes = [C.c_int64, f64p, i64p, C.c_int]
    L.bo_lu_factor.restype = C.c_int64
    L.bo_lu_solve.argtypes = [C.c_int64, f64p, i64p, f64p]
    L.bo_bicgstab.argtypes = [C.c_int64, f64p, f64p, f64p, C.c_double, C.c_int64, C.POINTER(C.c_double), C.c_int]
    L.bo_bicgstab.restype = C.c_int64
    for nm in ("bo_triangle_normals", "bo_vertex_normals"):
        getattr(L, nm).argtypes = [C.c_int64, f64p, C.c_int64, u32p, f64p]
    L.bo_tangent_gradients.argtypes = [C.c_int64, f64p, C.c_int64, u32p, f64p, f64p]
    L.bo_curvatures.argtypes = [C.c_int64, f64p, C.c_int64, u32p, f64p, f64p]
    L.bo_kappa.argtypes = [C.c_int64, f64p, C.c_int64, u32p, f64p]
    L.bo_components.argtypes = [C.c_int64, C.c_int64, u32p, i64p]
    L.bo_components.restype = C.c_int64
    L.bo_part_volumes.argtypes = [C.c_int64, f64p, C.c_int64, u32p, i64p, C.c_int64, f64p]
    L.bo_position_t.argtypes = [C.POINTER(Params), C.c_int64, f64p, C.c_int64, u32p, f64p, f64p, f64p]
    L.bo_pot_t_multi.argtypes = [C.POINTER(Params), C.c_int64, f64p, C.c_int64, u32p, f64p, C.c_double, f64p]
    L.bo_rk4_step.argtypes = [C.POINTER(Params), C.c_int64, f64p, C.c_int64, u32p, f64p, f64p,
                              C.POINTER(C.c_double), C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.bo_rk4_step.restype = C.c_int
    L.bo_euler_step.argtypes = [C.POINTER(Params), C.c_int64, f64p, C.c_int64, u32p, f64p, f64p,
                                C.POINTER(C.c_double), C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.bo_pin_boundary.argtypes = [C.c_int64, f64p, C.c_int64, u32p, C.c_void_p, C.c_void_p]
    L.bo_pin_boundary.restype = C.c_int64
    L.bo_drift_velocity.argtypes = [C.c_double]
    L.bo_drift_velocity.restype = C.c_double
    L.bo_euler_step.restype = C.c_int
    L.bo_compute_psi.argtypes = [C.POINTER(Params), C.c_int64, f64p, C.c_int64, u32p, f64p, f64p]
    L.bo_quadrature19.argtypes = [f64p] * 3
    L.bo_exterior_panel.argtypes = [f64p] * 6
    L.bo_exterior_pot.argtypes = [C.c_int64, f64p, C.c_int64, u32p, f64p, f64p, C.c_int64, f64p, f64p, C.c_int]
    L.bo_get_dt.argtypes = [C.POINTER(Params), C.c_int64, C.c_double, f64p, f64p]
    L.bo_get_dt.restype = C.c_double
    L.bo_nopenetration.argtypes = [C.c_int64, C.c_double, C.c_double, f64p, f64p, f64p]
    L.bo_lu_factor_blocked.argtypes = [C.c_int64, C.c_void_p, i64p, C.c_int, C.c_int64, C.c_int64]
    L.bo_lu_factor_blocked.restype = C.c_int64
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def quadrature():
    out = [np.zeros(7) for _ in range(5)]
    lib().bo_quadrature(*out)
    return {"gauss_x": out[0], "gauss_w": out[1], "tri_x": out[2], "tri_y": out[3], "tri_w": out[4]}


def quadrature19():
    out = [np.zeros(19) for _ in range(3)]
    lib().bo_quadrature19(*out)
    return {"tri_x": out[0], "tri_y": out[1], "tri_w": out[2]}


def exterior_panel(y, a, b, c):
    g, h = np.zeros(3), np.zeros(3)
    lib().bo_exterior_panel(_c(y, np.float64), _c(a, np.float64), _c(b, np.float64), _c(c, np.float64), g, h)
    return g, h


def exterior_pot(verts, tris, phi, psi, points, nthreads=0):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    pts = _c(points, np.float64).reshape(-1, 3)
    out = np.zeros(pts.shape[0])
    lib().bo_exterior_pot(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), _c(phi, np.float64),
                          _c(psi, np.float64), pts.shape[0], pts.reshape(-1), out, nthreads)
    return out


def panel(x, a, b, c, self_vertex=-1, mirror=False):
    g, h = np.zeros(3), np.zeros(3)
    lib().bo_panel(_c(x, np.float64), _c(a, np.float64), _c(b, np.float64), _c(c, np.float64),
                   int(self_vertex), int(bool(mirror)), g, h)
    return g, h


def assemble(verts, tris, mirror=False, nthreads=0):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    N, M = v.shape[0], t.shape[0]
    G = np.zeros((N, N), order="F")
    H = np.zeros((N, N), order="F")
    Gf, Hf = G.reshape(-1, order="F"), H.reshape(-1, order="F")
    # reshape of an F-ordered array with order F is a view
    lib().bo_assemble(N, v.reshape(-1), M, t.reshape(-1), int(bool(mirror)), Gf, Hf, nthreads)
    return G, H


def gemv(A, x, nthreads=0):
    A = np.asfortranarray(A, dtype=np.float64)
    y = np.zeros(A.shape[0])
    lib().bo_gemv(A.shape[0], A.reshape(-1, order="F"), _c(x, np.float64), y, nthreads)
    return y


def lu_factor(A, nthreads=0):
    LU = np.array(A, dtype=np.float64, order="F", copy=True)
    piv = np.zeros(LU.shape[0], dtype=np.int64)
    info = lib().bo_lu_factor(LU.shape[0], LU.reshape(-1, order="F"), piv, nthreads)
    return LU, piv, int(info)


def lu_factor_blocked(A, nthreads=0, nb=256, max_panels=0, in_place=False):
    """Cache-blocked GEPP (the CPU baseline's LU); in_place factors a Fortran-
    ordered float64 A without a copy."""
    LU = A if in_place else np.array(A, dtype=np.float64, order="F", copy=True)
    assert LU.flags.f_contiguous and LU.dtype == np.float64
    piv = np.zeros(LU.shape[0], dtype=np.int64)
    info = lib().bo_lu_factor_blocked(LU.shape[0], LU.ctypes.data, piv, nthreads, nb, max_panels)
    return LU, piv, int(info)


def lu_solve(LU, piv, b):
    x = np.array(b, dtype=np.float64, copy=True)
    lib().bo_lu_solve(LU.shape[0], np.asfortranarray(LU).reshape(-1, order="F"), _c(piv, np.int64), x)
    return x


def solve(A, b, nthreads=0):
    LU, piv, _ = lu_factor(A, nthreads)
    return lu_solve(LU, piv, b)


def bicgstab(A, b, tol=0.0, maxit=0, nthreads=0):
    A = np.asfortranarray(A, dtype=np.float64)
    x = np.zeros(A.shape[0])
    err = C.c_double(0.0)
    it = lib().bo_bicgstab(A.shape[0], A.reshape(-1, order="F"), _c(b, np.float64), x, tol, maxit, C.byref(err), nthreads)
    return x, int(it), err.value


def triangle_normals(verts, tris):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    out = np.zeros((t.shape[0], 3))
    lib().bo_triangle_normals(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), out.reshape(-1))
    return out


def vertex_normals(verts, tris):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    out = np.zeros((v.shape[0], 3))
    lib().bo_vertex_normals(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), out.reshape(-1))
    return out


def tangent_gradients(verts, tris, pot):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    out = np.zeros((t.shape[0], 3))
    lib().bo_tangent_gradients(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), _c(pot, np.float64), out.reshape(-1))
    return out


def curvatures(verts, tris):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    k, g = np.zeros(t.shape[0]), np.zeros(t.shape[0])
    lib().bo_curvatures(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), k, g)
    return k, g


def kappa(verts, tris):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    out = np.zeros(v.shape[0])
    lib().bo_kappa(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), out)
    return out


def components(nverts, tris):
    t = _c(tris, np.uint32)
    comp = np.zeros(nverts, dtype=np.int64)
    n = lib().bo_components(nverts, t.shape[0], t.reshape(-1), comp)
    return comp, int(n)


def part_volumes(verts, tris):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    comp, n = components(v.shape[0], t)
    vol = np.zeros(n)
    lib().bo_part_volumes(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), comp, n, vol)
    return vol


def make_params(physics=None, *, V_0=None, mirror=False, wall_guards=False, frozen_stage_time=True,
                solver="lu", nthreads=0, dp_balance=3.0, min_dt=-1.0, wall_eps=1e-2, verts=None, tris=None,
                n_pin=0):
    p = Params()
    ph = physics
    p.p_inf = getattr(ph, "p_inf", 1.0)
    p.epsilon = getattr(ph, "epsilon", 1.0)
    p.sigma = getattr(ph, "sigma", 0.0)
    p.gamma = getattr(ph, "gamma", 1.0)
    p.field_amp = getattr(ph, "field_amp", 0.0)
    p.field_freq = getattr(ph, "field_freq", 0.0)
    p.field_phase = getattr(ph, "field_phase", 0.0)
    if V_0 is None:
        V_0 = getattr(ph, "V_0", None)
    if V_0 is None:
        V_0 = float(part_volumes(verts, tris).sum())   # Simulation.hpp:42, total initial volume
    p.V_0 = V_0
    p.dp_balance, p.min_dt, p.wall_eps = dp_balance, min_dt, wall_eps
    p.mirror, p.wall_guards, p.frozen_stage_time = int(mirror), int(wall_guards), int(frozen_stage_time)
    p.solver = {"lu": 0, "bicgstab": 1}[solver]
    p.nthreads = nthreads
    p.n_pin = n_pin
    return p


def pin_boundary(verts, tris, phi=None, psi=None):
    """ColocSimPin::set_x_boundary: reordered copies (verts, tris, phi, psi) and the pinned count"""
    v = np.array(verts, dtype=np.float64, copy=True)
    t = np.array(tris, dtype=np.uint32, copy=True)
    ph = None if phi is None else np.array(phi, dtype=np.float64, copy=True)
    ps = None if psi is None else np.array(psi, dtype=np.float64, copy=True)
    n = lib().bo_pin_boundary(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1),
                              None if ph is None else ph.ctypes.data, None if ps is None else ps.ctypes.data)
    return v, t, ph, ps, int(n)


def drift_velocity(costheta):
    return lib().bo_drift_velocity(float(costheta))


def compute_psi(params, verts, tris, phi):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    psi = np.zeros(v.shape[0])
    lib().bo_compute_psi(C.byref(params), v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), _c(phi, np.float64), psi)
    return psi


def position_t(params, verts, tris, pot):
    v, t = _c(verts, np.float64), _c(tris, np.uint32)
    vel = np.zeros((v.shape[0], 3))
    psi = np.zeros(v.shape[0])
    lib().bo_position_t(C.byref(params), v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1),
                        _c(pot, np.float64), vel.reshape(-1), psi)
    return vel, psi


def pot_t_multi(params, verts, tris, vel, t):
    v, tr = _c(verts, np.float64), _c(tris, np.uint32)
    out = np.zeros(v.shape[0])
    lib().bo_pot_t_multi(C.byref(params), v.shape[0], v.reshape(-1), tr.shape[0], tr.reshape(-1),
                         _c(vel, np.float64).reshape(-1), float(t), out)
    return out


def get_dt(params, dp, vel, dpot):
    v = _c(vel, np.float64)
    return lib().bo_get_dt(C.byref(params), v.shape[0], float(dp), v.reshape(-1), _c(dpot, np.float64))


def nopenetration(eps, dt, x0, c):
    """(xf, c') of LinLinSim::nopenetration; c' is c with the rewritten c.x"""
    x = _c(x0, np.float64)
    cc = np.array(c, dtype=np.float64, copy=True)
    xf = np.zeros_like(x)
    lib().bo_nopenetration(x.shape[0], float(eps), float(dt), x.reshape(-1), cc.reshape(-1), xf.reshape(-1))
    return xf, cc


class Sim:
    """Oracle state stepper mirroring LinLinSim (x, phi, psi, time)."""

    def __init__(self, params, verts, tris, phi, library=None):
        self.lib = library if library is not None else lib()
        self.p = params
        self.x = np.array(verts, dtype=np.float64, copy=True)
        self.t = _c(tris, np.uint32)
        self.phi = np.array(phi, dtype=np.float64, copy=True)
        self.psi = np.zeros(self.x.shape[0])
        self.time = 0.0

    def rk4(self, dp, fixdt=True):
        tm = C.c_double(self.time)
        dt = C.c_double(0.0)
        rc = self.lib.bo_rk4_step(C.byref(self.p), self.x.shape[0], self.x.reshape(-1), self.t.shape[0],
                               self.t.reshape(-1), self.phi, self.psi, C.byref(tm), float(dp), int(fixdt), C.byref(dt))
        if rc != 0:
            raise IndexError("negative values")   # LinLinSim.cpp:157 throws out_of_range
        self.time = tm.value
        return dt.value

    def euler(self, dp, fixdt=True):
        tm = C.c_double(self.time)
        dt = C.c_double(0.0)
        self.lib.bo_euler_step(C.byref(self.p), self.x.shape[0], self.x.reshape(-1), self.t.shape[0],
                            self.t.reshape(-1), self.phi, self.psi, C.byref(tm), float(dp), int(fixdt), C.byref(dt))
        self.time = tm.value
        return dt.value
