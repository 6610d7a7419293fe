"""Python mirror of the reference's simulation interface over libbemhip's C-ABI.

The class `ColocSim` follows the method names and argument meaning of the
reference's Bem::ColocSim / LinLinSim / Simulation (Simulation.hpp:29-167,
LinLinSim.hpp:22-67): evolve_system_RK4, evolve_system, compute_psi,
set_phi/get_phi, get_psi, set_V_0, set_time/get_time, get_volume,
set_bcgstab, set_min_dt, set_dp_balance, position_t, pot_t_multi.  Errors are
raised like the reference raises them (IndexError for the wall-mode
"negative values" std::out_of_range) or as BemHipError with the library's
status code.

There is no CPU fallback: every call runs the gfx950 kernels of libbemhip.so,
and loading fails loudly when the library is missing.
"""

from __future__ import annotations

import ctypes as C
import math
import os
import re
import weakref

import numpy as np

# The library leaves the HIP runtime's hardware-queue count
# (GPU_MAX_HW_QUEUES, read once when the process first initialises HIP,
# default 4) to the application.  A batch of independent problems is solved
# on up to 8 solver lanes of two streams each, which then share those queues.

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbemhip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "bemhip.h")

(OK, ERR_ARG, ERR_BAD_MESH, ERR_NEG_X, ERR_SINGULAR, ERR_NOT_CONVERGED, ERR_NAN, ERR_OOM, ERR_HIP, ERR_NO_DEVICE,
 ERR_IO) = range(11)
MIRROR_X0 = 0x1
INDEPENDENT_PARTS = 0x2
SOLVERS = {"lu64": 0, "lu32_ir64": 1, "bicgstab64": 2, "lu16x3_ir64": 3, "lu16_ir64": 4}

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class BemHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bemhip status {code}: {msg}")
        self.code = code


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("residual", C.c_double),
                ("correction", C.c_double), ("ms_assemble", C.c_double), ("ms_factor", C.c_double),
                ("ms_solve", C.c_double), ("gemm_ms", C.c_double), ("gemm_flops", C.c_double),
                ("gemm_launches", C.c_int32), ("pad_", C.c_int32), ("panels", C.c_int32),
                ("panels_pivoted", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class TuneEntry(C.Structure):
    _fields_ = [("nb", C.c_int32), ("split", C.c_int32), ("lookahead", C.c_int32), ("reserve", C.c_int32),
                ("ms", C.c_double), ("chosen", C.c_int32), ("pad_", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class SimParams(C.Structure):
    _fields_ = [("p_inf", C.c_double), ("epsilon", C.c_double), ("sigma", C.c_double), ("gamma", C.c_double),
                ("V_0", C.c_double), ("field_amp", C.c_double), ("field_freq", C.c_double),
                ("field_phase", C.c_double), ("dp_balance", C.c_double), ("min_dt", C.c_double),
                ("wall_eps", C.c_double), ("tol", C.c_double), ("maxit", C.c_int32), ("solver", C.c_int32),
                ("flags", C.c_uint32), ("wall_guards", C.c_int32), ("frozen_stage_time", C.c_int32),
                ("pad_", C.c_int32)]


_lib = None


def declared_functions(header: str = HEADER):
    """Names of the functions declared by include/bemhip.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(bemhip_\w+)\s*\(", text, flags=re.M)))


def load(path: str = LIB_PATH):
    """Load libbemhip.so (no device work).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libbemhip.so not built at {path}; run __graft_entry__.build()")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.bemhip_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.bemhip_destroy.argtypes = [vp]
    L.bemhip_last_error.argtypes = [vp]
    L.bemhip_last_error.restype = C.c_char_p
    L.bemhip_quadrature.argtypes = [f64p]
    L.bemhip_quadrature.restype = None
    L.bemhip_assemble.argtypes = [vp, C.c_int64, f64p, C.c_int64, u32p, C.c_uint32, vp, vp, vp, vp]
    L.bemhip_solve.argtypes = [vp, C.c_int64, vp, f64p, f64p, C.c_int, C.c_double, C.c_int32, C.POINTER(SolveStats)]
    L.bemhip_compute_psi.argtypes = [vp, C.c_int64, f64p, C.c_int64, u32p, f64p, C.c_uint32, C.c_int,
                                     C.c_double, C.c_int32, f64p, C.POINTER(SolveStats)]
    L.bemhip_sim_default_params.argtypes = [C.POINTER(SimParams)]
    L.bemhip_sim_default_params.restype = None
    L.bemhip_sim_create.argtypes = [vp, C.POINTER(SimParams), C.c_int64, f64p, C.c_int64, u32p, C.POINTER(vp)]
    L.bemhip_sim_destroy.argtypes = [vp]
    L.bemhip_sim_set_phi.argtypes = [vp, f64p]
    L.bemhip_sim_set_psi.argtypes = [vp, f64p]
    L.bemhip_sim_set_V_0.argtypes = [vp, C.c_double]
    L.bemhip_sim_set_time.argtypes = [vp, C.c_double]
    L.bemhip_sim_set_solver.argtypes = [vp, C.c_int, C.c_double, C.c_int32]
    L.bemhip_sim_get_state.argtypes = [vp, vp, vp, vp, C.POINTER(C.c_double)]
    L.bemhip_sim_get_volume.argtypes = [vp, C.POINTER(C.c_double)]
    L.bemhip_sim_position_t.argtypes = [vp, vp, vp]
    L.bemhip_sim_pot_t_multi.argtypes = [vp, f64p, C.c_double, f64p]
    L.bemhip_sim_compute_psi.argtypes = [vp]
    L.bemhip_sim_evolve_rk4.argtypes = [vp, C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.bemhip_sim_evolve_euler.argtypes = [vp, C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.bemhip_sim_last_stats.argtypes = [vp, C.POINTER(SolveStats)]
    L.bemhip_sim_tune.argtypes = [vp, C.c_int32, C.POINTER(TuneEntry), C.POINTER(C.c_int32)]
    L.bemhip_sim_accum_stats.argtypes = [vp, C.POINTER(SolveStats), C.c_int]
    L.bemhip_sim_set_params.argtypes = [vp, C.POINTER(SimParams)]
    L.bemhip_sim_set_mesh.argtypes = [vp, C.c_int64, f64p, C.c_int64, u32p]
    L.bemhip_sim_get_parts.argtypes = [vp, C.POINTER(C.c_int32)]
    L.bemhip_sim_get_part_volumes.argtypes = [vp, f64p, C.c_int32]
    L.bemhip_sim_set_part_physics.argtypes = [vp, C.c_int32, vp, vp, vp, vp]
    L.bemhip_exterior_potential.argtypes = [vp, C.c_int64, f64p, C.c_int64, u32p, f64p, f64p, C.c_int64, f64p, f64p]
    L.bemhip_exterior_potential_noalloc.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, vp, vp, C.c_int64, vp, vp]
    L.bemhip_sim_exterior_pot.argtypes = [vp, C.c_int64, f64p, f64p]
    L.bemhip_sim_get_triangles.argtypes = [vp, u32p]
    L.bemhip_sim_export_ply.argtypes = [vp, C.c_char_p, C.c_int32]
    L.bemhip_pin_boundary.argtypes = [C.c_int64, f64p, C.c_int64, u32p, vp, vp, C.POINTER(C.c_int64)]
    L.bemhip_sim_create_pinned.argtypes = [vp, C.POINTER(SimParams), C.c_int64, f64p, C.c_int64, u32p, vp, vp,
                                           C.POINTER(vp), C.POINTER(C.c_int64)]
    L.bemhip_sim_set_pinned.argtypes = [vp, C.c_int64]
    L.bemhip_sim_get_pinned.argtypes = [vp, C.POINTER(C.c_int64)]
    L.bemhip_io_last_error.argtypes = []
    L.bemhip_io_last_error.restype = C.c_char_p
    L.bemhip_ply_write.argtypes = [C.c_char_p, C.c_int64, f64p, C.c_int64, u32p, vp, vp, C.c_int32]
    L.bemhip_ply_info.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]
    L.bemhip_ply_read.argtypes = [C.c_char_p, C.c_int64, f64p, C.c_int64, u32p, vp, vp]
    L.bemhip_sim_adjoint_radius.argtypes = [vp, C.c_int32, C.c_double, f64p, C.c_int32, C.POINTER(C.c_double),
                                            vp, vp, vp]
    _lib = L
    return L


def _io_check(L, rc):
    if rc != OK:
        raise BemHipError(rc, L.bemhip_io_last_error().decode())


def _vp(a):
    return None if a is None else a.ctypes.data


def write_ply(fname, verts, tris, phi=None, psi=None, precision=8):
    """export_ply_double / export_ply_float (MeshIO.cpp:263-335) and the one-value
    and bare-mesh variants (MeshIO.cpp:67-100, 191-261); no device needed."""
    L = load()
    v = np.ascontiguousarray(verts, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    ph = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
    ps = None if psi is None else np.ascontiguousarray(psi, dtype=np.float64)
    _io_check(L, L.bemhip_ply_write(os.fsencode(fname), v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1),
                                    _vp(ph), _vp(ps), int(precision)))


def read_ply(fname):
    """import_ply (MeshIO.cpp:701-903): (verts, tris, phi, psi); phi / psi are
    None when the file holds no such values."""
    L = load()
    N, M, nv, pr = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
    path = os.fsencode(fname)
    _io_check(L, L.bemhip_ply_info(path, C.byref(N), C.byref(M), C.byref(nv), C.byref(pr)))
    v = np.zeros((N.value, 3))
    t = np.zeros((M.value, 3), dtype=np.uint32)
    phi = np.zeros(N.value) if nv.value >= 1 else None
    psi = np.zeros(N.value) if nv.value >= 2 else None
    _io_check(L, L.bemhip_ply_read(path, N.value, v.reshape(-1), M.value, t.reshape(-1), _vp(phi), _vp(psi)))
    return v, t, phi, psi


def pin_boundary(verts, tris, phi=None, psi=None):
    """ColocSimPin::set_x_boundary (ColocSimPin.cpp:222-287) on copies:
    (verts, tris, phi, psi, n_pin); no device needed."""
    L = load()
    v = np.array(verts, dtype=np.float64, copy=True)
    t = np.array(tris, dtype=np.uint32, copy=True)
    ph = None if phi is None else np.array(phi, dtype=np.float64, copy=True)
    ps = None if psi is None else np.array(psi, dtype=np.float64, copy=True)
    n = C.c_int64()
    rc = L.bemhip_pin_boundary(v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1), _vp(ph), _vp(ps), C.byref(n))
    if rc != OK:
        raise BemHipError(rc, "pin_boundary: bad arguments or mesh")
    return v, t, ph, ps, n.value


def quadrature():
    out = np.zeros(35)
    load().bemhip_quadrature(out)
    return {"gauss_x": out[0:7], "gauss_w": out[7:14], "tri_x": out[14:21], "tri_y": out[21:28], "tri_w": out[28:35]}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _solver_id(s):
    return SOLVERS[s] if isinstance(s, str) else int(s)


class Context:
    """A device, stream and workspace (bemhip_create / bemhip_destroy)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.bemhip_create(device, C.byref(h))
        if rc != OK:
            raise BemHipError(rc, "bemhip_create failed (no gfx950 device?)")
        self.h = h
        self._sims = weakref.WeakSet()   # simulations living on this context

    def close(self):
        """Destroys the simulations still open on this context, then the context
        (a simulation must never outlive its context)."""
        if getattr(self, "h", None):
            for sim in list(getattr(self, "_sims", ())):
                sim.close()
            self.L.bemhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != OK:
            raise BemHipError(rc, self.L.bemhip_last_error(self.h).decode())

    # ColocSim::assemble_matrices (ColocSim.cpp:21-110)
    def assemble(self, verts, tris, mirror=False, want_H=True, phi=None, want_G=True):
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        N, M = v.shape[0], t.shape[0]
        G = np.zeros((N, N), order="F") if want_G else None
        H = np.zeros((N, N), order="F") if want_H else None
        ph = np.ascontiguousarray(phi, dtype=np.float64) if phi is not None else None
        hphi = np.zeros(N) if phi is not None else None
        self.check(self.L.bemhip_assemble(self.h, N, v.reshape(-1), M, t.reshape(-1),
                                          MIRROR_X0 if mirror else 0, _ptr(G), _ptr(H), _ptr(ph), _ptr(hphi)))
        return G, H, hphi

    # Simulation::solve_system (Simulation.cpp:40-71)
    def solve(self, G, b, solver="lu64", tol=0.0, maxit=0):
        Gf = np.asfortranarray(G, dtype=np.float64)
        x = np.zeros(Gf.shape[0])
        st = SolveStats()
        self.check(self.L.bemhip_solve(self.h, Gf.shape[0], _ptr(Gf), np.ascontiguousarray(b, dtype=np.float64),
                                       x, _solver_id(solver), tol, maxit, C.byref(st)))
        return x, st.as_dict()

    # Simulation::compute_psi (Simulation.hpp:73-77)
    def compute_psi(self, verts, tris, phi, solver="lu64", mirror=False, tol=0.0, maxit=0, independent_parts=False):
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        psi = np.zeros(v.shape[0])
        st = SolveStats()
        flags = (MIRROR_X0 if mirror else 0) | (INDEPENDENT_PARTS if independent_parts else 0)
        self.check(self.L.bemhip_compute_psi(self.h, v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1),
                                             np.ascontiguousarray(phi, dtype=np.float64),
                                             flags, _solver_id(solver), tol, maxit, psi, C.byref(st)))
        return psi, st.as_dict()

    # no-allocation variant (no reference counterpart): device pointers (ints)
    # of caller-owned buffers; xyz N x 3, tri M x 3 uint32, phi / psi N,
    # points P x 3, out P
    def exterior_potential_noalloc(self, N, d_xyz, M, d_tri, d_phi, d_psi, P, d_points, d_out):
        self.check(self.L.bemhip_exterior_potential_noalloc(self.h, N, d_xyz, M, d_tri, d_phi, d_psi, P, d_points,
                                                            d_out))

    # compute_exterior_pot (LinLinSim.cpp:440-477)
    def exterior_potential(self, verts, tris, phi, psi, points):
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(pts.shape[0])
        self.check(self.L.bemhip_exterior_potential(self.h, v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1),
                                                    np.ascontiguousarray(phi, dtype=np.float64),
                                                    np.ascontiguousarray(psi, dtype=np.float64),
                                                    pts.shape[0], pts.reshape(-1), out))
        return out


class ColocSim:
    """Device-resident drop-in for Bem::ColocSim (ColocSim.hpp:16-22)."""

    def __init__(self, ctx: Context, verts, tris, p_inf=1.0, epsilon=1.0, sigma=0.0, gamma=1.0,
                 field_amp=0.0, field_freq=0.0, field_phase=0.0, mirror=False, wall_guards=None,
                 frozen_stage_time=True, solver="bicgstab64", tol=0.0, maxit=0, V_0=None, independent_parts=False):
        self.ctx = ctx
        L = ctx.L
        p = SimParams()
        L.bemhip_sim_default_params(C.byref(p))
        p.p_inf, p.epsilon, p.sigma, p.gamma = p_inf, epsilon, sigma, gamma
        p.field_amp, p.field_freq, p.field_phase = field_amp, field_freq, field_phase
        p.flags = (MIRROR_X0 if mirror else 0) | (INDEPENDENT_PARTS if independent_parts else 0)
        p.wall_guards = int(mirror if wall_guards is None else wall_guards)
        p.frozen_stage_time = int(frozen_stage_time)
        p.solver, p.tol, p.maxit = _solver_id(solver), tol, maxit
        if V_0 is not None:
            p.V_0 = V_0
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        self.N, self.M = v.shape[0], t.shape[0]
        self.h = self._create(L, p, v, t)
        ctx._sims.add(self)

    def _create(self, L, p, v, t):
        h = C.c_void_p()
        self.ctx.check(L.bemhip_sim_create(self.ctx.h, C.byref(p), self.N, v.reshape(-1), self.M, t.reshape(-1),
                                           C.byref(h)))
        return h

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):   # the context destroys its simulations first when it closes
                self.ctx.L.bemhip_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc == ERR_NEG_X:
            raise IndexError("negative values")   # LinLinSim.cpp:157 throws std::out_of_range
        self.ctx.check(rc)

    # setters / getters (Simulation.hpp:92-167)
    def set_phi(self, phi):
        phi = np.ascontiguousarray(np.broadcast_to(np.asarray(phi, dtype=np.float64), (self.N,)))
        self._check(self.ctx.L.bemhip_sim_set_phi(self.h, phi))

    def set_psi(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        if psi.shape != (self.N,):
            raise ValueError(f"psi must have shape ({self.N},), got {psi.shape}")
        self._check(self.ctx.L.bemhip_sim_set_psi(self.h, psi))

    def set_V_0(self, v):
        self._check(self.ctx.L.bemhip_sim_set_V_0(self.h, float(v)))

    def set_time(self, t):
        self._check(self.ctx.L.bemhip_sim_set_time(self.h, float(t)))

    def set_solver(self, solver, tol=0.0, maxit=0):
        self._check(self.ctx.L.bemhip_sim_set_solver(self.h, _solver_id(solver), tol, maxit))

    def set_bcgstab(self, value: bool):   # Simulation.hpp:153-155
        self.set_solver("bicgstab64" if value else "lu64")

    def set_mesh(self, verts, tris):
        """New vertex positions / connectivity (remeshing); phi and psi are reset
        to zero, time and V_0 are kept."""
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        self._check(self.ctx.L.bemhip_sim_set_mesh(self.h, v.shape[0], v.reshape(-1), t.shape[0], t.reshape(-1)))
        self.N, self.M = v.shape[0], t.shape[0]

    def state(self):
        x = np.zeros((self.N, 3))
        phi = np.zeros(self.N)
        psi = np.zeros(self.N)
        t = C.c_double()
        self._check(self.ctx.L.bemhip_sim_get_state(self.h, _ptr(x), _ptr(phi), _ptr(psi), C.byref(t)))
        return x, phi, psi, t.value

    def get_vertices(self):
        return self.state()[0]

    def get_phi(self):
        return self.state()[1]

    def get_psi(self):
        return self.state()[2]

    def get_time(self):
        return self.state()[3]

    def get_volume(self):
        v = C.c_double()
        self._check(self.ctx.L.bemhip_sim_get_volume(self.h, C.byref(v)))
        return v.value

    # loose parts (Mesh.cpp:382-440) and per-part physics of a batch
    def nparts(self):
        n = C.c_int32()
        self._check(self.ctx.L.bemhip_sim_get_parts(self.h, C.byref(n)))
        return n.value

    def part_volumes(self):
        n = self.nparts()
        out = np.zeros(n)
        self._check(self.ctx.L.bemhip_sim_get_part_volumes(self.h, out, n))
        return out

    def adjoint_radius(self, nsteps, dt, target, every=10, want_grad=True):
        """Loss sum_n (R_n - target_n)^2 over `nsteps` fixed RK4 steps from the
        current state, the radii R_n, and (when want_grad) dL/dx0 (N x 3) and
        dL/dphi0 (N); the simulation's state is left unchanged."""
        tgt = np.ascontiguousarray(target, dtype=np.float64)
        assert tgt.shape == (nsteps,)
        loss = C.c_double()
        R = np.zeros(nsteps)
        gx = np.zeros((self.N, 3)) if want_grad else None
        gp = np.zeros(self.N) if want_grad else None
        self._check(self.ctx.L.bemhip_sim_adjoint_radius(self.h, nsteps, dt, tgt, every, C.byref(loss), _ptr(R),
                                                         _ptr(gx), _ptr(gp)))
        return loss.value, R, gx, gp

    def set_part_physics(self, V_0=None, field_amp=None, field_freq=None, field_phase=None):
        """Per-part V_0 and forcing p_field = amp sin(2 pi freq t + phase); a
        parameter given as None is left as it is."""
        n = self.nparts()
        arrs = [None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
                for a in (V_0, field_amp, field_freq, field_phase)]
        self._check(self.ctx.L.bemhip_sim_set_part_physics(self.h, n, *[_ptr(a) for a in arrs]))

    # hot path
    def position_t(self):
        vel = np.zeros((self.N, 3))
        psi = np.zeros(self.N)
        self._check(self.ctx.L.bemhip_sim_position_t(self.h, _ptr(vel), _ptr(psi)))
        return vel, psi

    def pot_t_multi(self, vel, t):
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        if vel.size != 3 * self.N:
            raise ValueError(f"vel must hold {self.N} x 3 values, got shape {vel.shape}")
        out = np.zeros(self.N)
        self._check(self.ctx.L.bemhip_sim_pot_t_multi(self.h, vel.reshape(-1), float(t), out))
        return out

    def compute_psi(self):
        self._check(self.ctx.L.bemhip_sim_compute_psi(self.h))
        return self.get_psi()

    # Simulation::export_mesh (Simulation.cpp:94-96) writes floats; precision=8
    # is export_ply_double (MeshIO.cpp:300-335), which restores the state exactly
    def export_mesh(self, fname, precision=4):
        self._check(self.ctx.L.bemhip_sim_export_ply(self.h, os.fsencode(fname), int(precision)))

    def get_triangles(self):
        t = np.zeros((self.M, 3), dtype=np.uint32)
        self._check(self.ctx.L.bemhip_sim_get_triangles(self.h, t.reshape(-1)))
        return t

    @property
    def N_pin(self):   # ColocSimPin.hpp:57
        n = C.c_int64(0)
        self._check(self.ctx.L.bemhip_sim_get_pinned(self.h, C.byref(n)))
        return n.value

    @N_pin.setter
    def N_pin(self, value):
        self._check(self.ctx.L.bemhip_sim_set_pinned(self.h, int(value)))

    # LinLinSim::exterior_pot (LinLinSim.cpp:478-483)
    def exterior_pot(self, points):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(pts.shape[0])
        self._check(self.ctx.L.bemhip_sim_exterior_pot(self.h, pts.shape[0], pts.reshape(-1), out))
        return out

    def evolve_system_RK4(self, dp, fixdt=False):
        dt = C.c_double()
        self._check(self.ctx.L.bemhip_sim_evolve_rk4(self.h, float(dp), int(fixdt), C.byref(dt)))
        return dt.value

    def evolve_system(self, dp, fixdt=False):
        dt = C.c_double()
        self._check(self.ctx.L.bemhip_sim_evolve_euler(self.h, float(dp), int(fixdt), C.byref(dt)))
        return dt.value

    def last_stats(self):
        st = SolveStats()
        self._check(self.ctx.L.bemhip_sim_last_stats(self.h, C.byref(st)))
        return st.as_dict()

    # solver tuning (no reference counterpart): every candidate setting timed
    # on the current state, the fastest kept on the context
    def tune(self):
        n = C.c_int32(0)
        buf = (TuneEntry * 64)()
        self._check(self.ctx.L.bemhip_sim_tune(self.h, 64, buf, C.byref(n)))
        return [buf[i].as_dict() for i in range(min(64, n.value))]

    def accum_stats(self, reset=False):
        st = SolveStats()
        self._check(self.ctx.L.bemhip_sim_accum_stats(self.h, C.byref(st), int(reset)))
        return st.as_dict()


class ColocSimPin(ColocSim):
    """Drop-in for Bem::ColocSimPin (ColocSimPin.hpp:13-68): a bubble on the wall
    x = 0 whose contact-line vertices (|x| < 1e-5) are pinned; the vertices are
    reordered on construction (pinned last), as in the reference.  Wall mode
    (mirror) is on by default, as the reference compiles it (LinLinSim.hpp:13)."""

    def __init__(self, ctx: Context, verts, tris, *args, phi=None, psi=None, mirror=True, **kw):
        self._phi0 = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        self._psi0 = None if psi is None else np.ascontiguousarray(psi, dtype=np.float64)
        super().__init__(ctx, verts, tris, *args, mirror=mirror, **kw)

    def _create(self, L, p, v, t):
        h = C.c_void_p()
        n = C.c_int64()
        self.ctx.check(L.bemhip_sim_create_pinned(self.ctx.h, C.byref(p), self.N, v.reshape(-1), self.M,
                                                  t.reshape(-1), _vp(self._phi0), _vp(self._psi0), C.byref(h),
                                                  C.byref(n)))
        return h
