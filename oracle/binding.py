"""ctypes binding of oracle/liboracle.so (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        i, d = C.c_int, C.c_double
        lib.orc_Mss.argtypes = [_dp, _dp, _dp, i, i, i, d, d, i]
        lib.orc_Mbb.argtypes = [_dp, _dp, _dp, i, i, i, d, d, i]
        lib.orc_Msb.argtypes = [_dp, _dp, _dp, i, _dp, i, i, i, d, d, d, i]
        lib.orc_Mbs.argtypes = [_dp, _dp, _dp, i, _dp, i, i, i, d, d, d, i]
        lib.orc_barrier.argtypes = [_dp, _dp, _dp, i, _dp, i, i, d, d, d, i]
        lib.orc_ss_block.argtypes = [_dp, _dp, _dp, d, d]
        for fn in ("orc_lie_exp", "orc_lie_rotmat"):
            getattr(lib, fn).argtypes = [_dp, _dp]
            getattr(lib, fn).restype = None
        for fn in ("orc_lie_midpoint", "orc_lie_dexp", "orc_lie_dexpinv",
                   "orc_lie_dexpinv_transpose"):
            getattr(lib, fn).argtypes = [_dp, _dp, _dp]
            getattr(lib, fn).restype = None
        lib.orc_invert.argtypes = [_dp, i]
        lib.orc_Mss_wall.argtypes = [_dp, _dp, _dp, i, i, i, d, d, i]
        lib.orc_Mss_wall.restype = None
        lib.orc_seed_uniform.argtypes = [i, i, i, _dp]
        lib.orc_seed_band.argtypes = [i, i, d, _dp]
        lib.orc_seed_band.restype = None
        for f in (lib.orc_Mss, lib.orc_Mbb, lib.orc_Msb, lib.orc_Mbs,
                  lib.orc_barrier, lib.orc_ss_block):
            f.restype = None
        _lib = lib
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _c(a, cols):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, cols))


def nthreads_default():
    return max(1, min(16, os.cpu_count() or 1))


def Mss(xs, fs, a, mu=1.0, rows=None, nthreads=None):
    xs, fs = _c(xs, 3), _c(fs, 6)
    n = xs.shape[0]
    i0, i1 = rows if rows else (0, n)
    V = np.zeros((n, 6))
    load().orc_Mss(_p(V), _p(fs), _p(xs), n, i0, i1, a, mu, nthreads or nthreads_default())
    return V


def Mss_wall(xs, fs, a, mu=1.0, rows=None, nthreads=None):
    """Segment-segment mobility above the wall z = 0 (orc_Mss_wall)."""
    xs, fs = _c(xs, 3), _c(fs, 6)
    n = xs.shape[0]
    i0, i1 = rows if rows else (0, n)
    V = np.zeros((n, 6))
    load().orc_Mss_wall(_p(V), _p(fs), _p(xs), n, i0, i1, a, mu, nthreads or nthreads_default())
    return V


def Mbb(xb, fb, a, mu=1.0, rows=None, nthreads=None):
    xb, fb = _c(xb, 3), _c(fb, 3)
    n = xb.shape[0]
    i0, i1 = rows if rows else (0, n)
    V = np.zeros((n, 3))
    load().orc_Mbb(_p(V), _p(fb), _p(xb), n, i0, i1, a, mu, nthreads or nthreads_default())
    return V


def Msb(xs, xb, fb, a_s, a_b, mu=1.0, V=None, rows=None, nthreads=None):
    """V_s += M_sb F_b (V defaults to zeros)."""
    xs, xb, fb = _c(xs, 3), _c(xb, 3), _c(fb, 3)
    ns, nb = xs.shape[0], xb.shape[0]
    V = np.zeros((ns, 6)) if V is None else np.ascontiguousarray(V, dtype=np.float64).copy()
    i0, i1 = rows if rows else (0, ns)
    load().orc_Msb(_p(V), _p(fb), _p(xs), ns, _p(xb), nb, i0, i1, a_s, a_b, mu,
                   nthreads or nthreads_default())
    return V


def Mbs(xb, xs, fs, a_s, a_b, mu=1.0, rows=None, nthreads=None):
    xb, xs, fs = _c(xb, 3), _c(xs, 3), _c(fs, 6)
    ns, nb = xs.shape[0], xb.shape[0]
    V = np.zeros((nb, 3))
    i0, i1 = rows if rows else (0, nb)
    load().orc_Mbs(_p(V), _p(fs), _p(xb), nb, _p(xs), ns, i0, i1, a_s, a_b, mu,
                   nthreads or nthreads_default())
    return V


def grand(xs, xb, fs, fb, a_s, a_b, mu=1.0, seg_rows=None, blob_rows=None, nthreads=None):
    """(V_s, V_b) = all four reference blocks."""
    vs = Mss(xs, fs, a_s, mu, rows=seg_rows, nthreads=nthreads)
    vs = Msb(xs, xb, fb, a_s, a_b, mu, V=vs, rows=seg_rows, nthreads=nthreads)
    vb = Mbs(xb, xs, fs, a_s, a_b, mu, rows=blob_rows, nthreads=nthreads)
    vb += Mbb(xb, fb, a_b, mu, rows=blob_rows, nthreads=nthreads)
    return vs, vb


def barrier(xs, xb, fs, nseg, a_s, a_b, K, nthreads=None):
    """Returns (f_segs + barrier, f_blobs_repulsion)."""
    xs, xb = _c(xs, 3), _c(xb, 3)
    fs = _c(fs, 6).copy()
    frep = np.zeros((xb.shape[0], 3))
    load().orc_barrier(_p(fs), _p(frep), _p(xs), xs.shape[0], _p(xb), xb.shape[0],
                       nseg, a_s, a_b, K, nthreads or nthreads_default())
    return fs, frep


def ss_block(xi, xj, a, mu=1.0):
    out = np.zeros(36)
    load().orc_ss_block(_p(out), _p(np.ascontiguousarray(xi, float)),
                        _p(np.ascontiguousarray(xj, float)), a, mu)
    return out.reshape(6, 6)


def seed_uniform(n, order=1, max_it=200000):
    """Reference uniform seeding (seed.cpp); returns (points [n, 3], iterations)."""
    X = np.zeros(3 * n)
    it = load().orc_seed_uniform(int(n), int(order), int(max_it), _p(X))
    return X.reshape(-1, 3), it


def seed_band(n, platy, r_over_l):
    X = np.zeros(3 * n)
    load().orc_seed_band(int(n), int(bool(platy)), float(r_over_l), _p(X))
    return X.reshape(-1, 3)


# ---------------------------------------------------------------- time step

class OrcSimParams(C.Structure):
    _fields_ = [
        ("n_swimmers", C.c_int), ("nfil", C.POINTER(C.c_int)), ("nblob", C.POINTER(C.c_int)),
        ("nseg", C.c_int), ("blob_refs", _dp), ("fil_refs", _dp), ("body_x", _dp),
        ("fil_phase", _dp), ("mu", C.c_double), ("a_seg", C.c_double), ("a_blob", C.c_double),
        ("kb", C.c_double), ("kt", C.c_double), ("dt", C.c_double), ("tol", C.c_double),
        ("jacobian_confidence", C.c_double), ("gmres_tol", C.c_double),
        ("max_broyden_iter", C.c_int), ("gmres_max_iter", C.c_int),
        ("num_euler_steps", C.c_int), ("barrier_K", C.c_double),
        ("curvature_amp", C.c_double), ("curvature_period", C.c_double),
        ("sphere_radius", C.c_double), ("sphere_radii", _dp),
        ("end_force", C.c_double), ("perturbation", _dp), ("interact", C.c_int),
    ]


def _sim_lib():
    lib = load()
    if not getattr(lib, "_sim_ready", False):
        vp = C.c_void_p
        lib.orc_sim_create.restype = vp
        lib.orc_sim_create.argtypes = [C.POINTER(OrcSimParams), C.c_int]
        lib.orc_sim_destroy.argtypes = [vp]
        lib.orc_sim_destroy.restype = None
        lib.orc_sim_make_body_reference_matrices.argtypes = [vp]
        lib.orc_sim_step.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        for n in ("orc_sim_num_segments", "orc_sim_num_blobs", "orc_sim_nbroy"):
            getattr(lib, n).argtypes = [vp]
        lib.orc_sim_get_state.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        lib.orc_sim_get_state.restype = None
        lib.orc_sim_get_mobility.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        lib.orc_sim_get_mobility.restype = None
        lib.orc_sim_get_broyden.argtypes = [vp, _dp]
        lib.orc_sim_get_broyden.restype = None
        lib.orc_sim_test_compute_velocities.argtypes = [vp]
        lib.orc_sim_test_forces.argtypes = [vp]
        lib.orc_sim_test_forces.restype = None
        lib.orc_sim_test_prepare_jacobian.argtypes = [vp, C.c_int]
        lib.orc_sim_test_filament_jacobian.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        lib.orc_sim_test_jacobian_inv_mult.argtypes = [vp, _dp, _dp]
        lib.orc_sim_test_jacobian_inv_mult.restype = None
        lib.orc_sim_test_set_driving.argtypes = [vp, C.c_int]
        lib.orc_sim_test_set_driving.restype = None
        lib.orc_sim_test_errors.argtypes = [vp, C.c_int, _dp]
        lib.orc_sim_test_initial_guess.argtypes = [vp, C.c_int]
        lib.orc_sim_test_initial_guess.restype = None
        lib.orc_sim_backup_size.argtypes = [vp]
        lib.orc_sim_get_backup.argtypes = [vp, _dp]
        lib.orc_sim_get_backup.restype = None
        lib.orc_sim_set_backup.argtypes = [vp, _dp]
        lib.orc_sim_set_backup.restype = None
        lib.orc_test_filament_forces.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                                 _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                                 _dp, _dp]
        lib.orc_test_filament_forces.restype = None
        lib.orc_sim_test_set_kinematics.argtypes = [vp, _dp, _dp, _dp, _dp]
        lib.orc_sim_test_set_kinematics.restype = None
        lib.orc_sim_test_broyden_begin.argtypes = [vp, C.c_int]
        lib.orc_sim_test_broyden_iterate.argtypes = [vp, C.c_int, _dp, _dp, _dp]
        lib.orc_sim_test_broyden_columns.argtypes = [vp, _dp, _dp]
        lib._sim_ready = True
    return lib


class Sim:
    """Oracle time stepper for a geometry.SimConfig."""

    def __init__(self, cfg, nthreads=None):
        self.lib = _sim_lib()
        a = cfg.arrays()
        self._keep = a
        g0 = cfg.swimmers[0]
        self.p = OrcSimParams(
            len(a["nfil"]), a["nfil"].ctypes.data_as(C.POINTER(C.c_int)),
            a["nblob"].ctypes.data_as(C.POINTER(C.c_int)), cfg.nseg,
            _p(a["blob_refs"]), _p(a["fil_refs"]), _p(a["body_x"]), _p(a["phases"]),
            g0.mu, g0.a_seg, g0.a_blob, cfg.kb, cfg.kt, cfg.dt, cfg.tol,
            cfg.jacobian_confidence, cfg.gmres_tol, cfg.max_broyden_iter, cfg.gmres_max_iter,
            cfg.num_euler_steps, cfg.barrier_K, cfg.curvature_amp, cfg.curvature_period,
            float(a["radius"][0]), _p(a["radius"]), cfg.end_force,
            _p(a["perturbation"]) if a["perturbation"] is not None else None,
            int(bool(cfg.interact)))
        self.h = self.lib.orc_sim_create(C.byref(self.p), nthreads or nthreads_default())
        assert self.h, "oracle refused the configuration"
        self.Ns = self.lib.orc_sim_num_segments(self.h)
        self.Nb = self.lib.orc_sim_num_blobs(self.h)
        self.nbroy = self.lib.orc_sim_nbroy(self.h)
        self.nsw = len(a["nfil"])
        self.nfil_total = int(a["nfil"].sum())
        self.nseg = cfg.nseg

    def close(self):
        if self.h:
            self.lib.orc_sim_destroy(self.h)
            self.h = None

    def backup_values(self):
        out = np.zeros(self.lib.orc_sim_backup_size(self.h))
        self.lib.orc_sim_get_backup(self.h, _p(out))
        return out

    def set_backup(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.lib.orc_sim_backup_size(self.h)
        self.lib.orc_sim_set_backup(self.h, _p(v))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def make_body_reference_matrices(self):
        rc = self.lib.orc_sim_make_body_reference_matrices(self.h)
        assert rc == 0, rc

    def step(self, nt):
        bi, gi = C.c_int(0), C.c_int(0)
        rc = self.lib.orc_sim_step(self.h, nt, C.byref(bi), C.byref(gi))
        return rc, bi.value, gi.value

    def state(self):
        S = self.nseg
        x = np.zeros(3 * self.Ns); q = np.zeros(4 * self.Ns); u = np.zeros(3 * self.Ns)
        lam = np.zeros(self.nfil_total * (3 * (S - 1) + 6)); body = np.zeros(10 * self.nsw)
        self.lib.orc_sim_get_state(self.h, _p(x), _p(q), _p(u), _p(lam), _p(body))
        return dict(x=x.reshape(-1, 3), q=q.reshape(-1, 4), u=u.reshape(-1, 3),
                    lam=lam.reshape(self.nfil_total, -1), body=body.reshape(-1, 10))

    def mobility(self):
        vs = np.zeros(6 * self.Ns); vb = np.zeros(6 * self.nsw); fb = np.zeros(3 * self.Nb)
        fs = np.zeros(6 * self.Ns); ref = np.zeros(72 * self.nsw)
        self.lib.orc_sim_get_mobility(self.h, _p(vs), _p(vb), _p(fb), _p(fs), _p(ref))
        ref = ref.reshape(self.nsw, 2, 6, 6)
        return dict(v_segs=vs.reshape(-1, 6), v_bodies=vb.reshape(-1, 6),
                    f_blobs=fb.reshape(-1, 3), f_segs=fs.reshape(-1, 6),
                    Nref=ref[:, 0], KTKinv=ref[:, 1])

    def broyden_error(self):
        e = np.zeros(self.nbroy)
        self.lib.orc_sim_get_broyden(self.h, _p(e))
        return e

    # pieces of one step, for stage-by-stage parity
    def initial_guess(self, nt):
        self.lib.orc_sim_test_initial_guess(self.h, nt)

    def forces(self):
        self.lib.orc_sim_test_forces(self.h)

    def compute_velocities(self):
        return self.lib.orc_sim_test_compute_velocities(self.h)

    def errors(self, nt):
        e = np.zeros(self.nbroy)
        big = self.lib.orc_sim_test_errors(self.h, nt, _p(e))
        return e, bool(big)

    def prepare_jacobian(self, nt):
        rc = self.lib.orc_sim_test_prepare_jacobian(self.h, nt)
        assert rc == 0, rc

    def jacobian_inv_mult(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros_like(v)
        self.lib.orc_sim_test_jacobian_inv_mult(self.h, _p(v), _p(out))
        return out

    def set_kinematics(self, seg, v_segs, body, v_bodies):
        """seg: (Ns, 15) = x, xm1, xm2, u, um1; body: (nsw, 15); v_segs (Ns, 6);
        v_bodies (nsw, 6)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                for a in (seg, v_segs, body, v_bodies)]
        assert arrs[0].size == 15 * self.Ns and arrs[1].size == 6 * self.Ns
        assert arrs[2].size == 15 * self.nsw and arrs[3].size == 6 * self.nsw
        self.lib.orc_sim_test_set_kinematics(self.h, *[_p(a) for a in arrs])

    def broyden_begin(self, nt):
        rc = self.lib.orc_sim_test_broyden_begin(self.h, nt)
        assert rc >= 0, rc
        return bool(rc)

    def broyden_iterate(self, nt):
        dx, eo, en = (np.zeros(self.nbroy) for _ in range(3))
        big = self.lib.orc_sim_test_broyden_iterate(self.h, nt, _p(dx), _p(eo), _p(en))
        return dx, eo, en, bool(big)

    def broyden_columns(self, max_iter):
        Cm = np.zeros(self.nbroy * max_iter); Dm = np.zeros(self.nbroy * max_iter)
        it = self.lib.orc_sim_test_broyden_columns(self.h, _p(Cm), _p(Dm))
        return Cm.reshape(self.nbroy, max_iter), Dm.reshape(self.nbroy, max_iter), it

    def filament_jacobian(self, fil, nt):
        n = 6 * self.nseg
        J = np.zeros(n * n); Ji = np.zeros(n * n)
        rc = self.lib.orc_sim_test_filament_jacobian(self.h, fil, nt, _p(J), _p(Ji))
        assert rc == 0, rc
        return J.reshape(n, n), Ji.reshape(n, n)


# ---------------------------------------------------------------- Lie helpers

def _v(a, n):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n))


def lie_exp(u):
    q = np.zeros(4)
    load().orc_lie_exp(_p(_v(u, 3)), _p(q))
    return q


def lie_midpoint(q1, q2):
    q = np.zeros(4)
    load().orc_lie_midpoint(_p(_v(q1, 4)), _p(_v(q2, 4)), _p(q))
    return q


def lie_rotmat(q):
    R = np.zeros(9)
    load().orc_lie_rotmat(_p(_v(q, 4)), _p(R))
    return R.reshape(3, 3)


def _uv(fn, u, v):
    out = np.zeros(3)
    getattr(load(), fn)(_p(_v(u, 3)), _p(_v(v, 3)), _p(out))
    return out


def dexp(u, v):
    return _uv("orc_lie_dexp", u, v)


def dexpinv(u, v):
    return _uv("orc_lie_dexpinv", u, v)


def dexpinv_transpose(u, v):
    return _uv("orc_lie_dexpinv_transpose", u, v)


def filament_forces(q, lam, teth, clamp, u0, st, kb, kt, DL, end_force=0.0, pert=None):
    """Internal forces/torques (S, 6) of one filament in a given state."""
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    S = q.size // 4
    f = np.zeros(6 * S)
    args = [_v(lam, 3 * (S - 1)), _v(teth, 3), _v(clamp, 3), _v(u0, 3), _v(st, 3)]
    lam_, teth_, clamp_, u0_, st_ = args
    pp = _p(_v(pert, 6)) if pert is not None else None
    load().orc_test_filament_forces(S, DL, kb, kt, _p(q), _p(u0_), _p(lam_), _p(teth_),
                                    _p(clamp_), _p(st_), end_force, pp, _p(f))
    return f.reshape(S, 6)


def invert(A):
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    rc = load().orc_invert(_p(A), A.shape[0])
    if rc != 0:
        raise np.linalg.LinAlgError("singular")
    return A


# ---- prescribed cilia (oracle_cilia.h) -------------------------------------

class OrcCiliaParams(C.Structure):
    _fields_ = [
        ("n_swimmers", C.c_int), ("nfil", C.POINTER(C.c_int)), ("nblob", C.POINTER(C.c_int)),
        ("nseg", C.c_int), ("interact", C.c_int), ("blob_refs", _dp), ("fil_refs", _dp),
        ("body_x", _dp), ("phase", _dp), ("omega0", _dp), ("r_over_l", C.c_double),
        ("mu", C.c_double), ("a_seg", C.c_double), ("a_blob", C.c_double), ("dt", C.c_double),
        ("beat", C.c_int), ("amplitude", C.c_double), ("recovery_window", C.c_double),
        ("effective_stroke", C.c_double), ("zero_velocity_avoidance", C.c_double),
        ("dynamic_phase", C.c_int), ("fixed_body", C.c_int), ("body_velocity", _dp),
        ("gen_force_refs", _dp), ("n_gen_force_refs", C.c_int), ("gmres_tol", C.c_double),
        ("gmres_max_iter", C.c_int), ("wall", C.c_int),
    ]


def _cilia_lib():
    lib = load()
    if getattr(lib, "_cilia_ready", False):
        return lib
    i, v, cp = C.c_int, C.c_void_p, C.POINTER(C.c_int)
    sig = [
        ("orc_cilia_create", v, [C.POINTER(OrcCiliaParams), i]),
        ("orc_cilia_destroy", None, [v]),
        ("orc_cilia_system_size", i, [v]),
        ("orc_cilia_step", i, [v, i, cp]),
        ("orc_cilia_initial_guess", None, [v, i]),
        ("orc_cilia_compute_velocities", i, [v, i, cp]),
        ("orc_cilia_integrate_bodies", None, [v]),
        ("orc_cilia_end_of_step", None, [v, i]),
        ("orc_cilia_rhs", None, [v, _dp]),
        ("orc_cilia_apply_operator", None, [v, _dp, _dp]),
        ("orc_cilia_apply_preconditioner", None, [v, _dp, _dp]),
        ("orc_cilia_precon_matrices", None, [v, _dp]),
        ("orc_cilia_get_state", None, [v, _dp, _dp, _dp, _dp, _dp, _dp]),
        ("orc_cilia_get_solution", None, [v, _dp, _dp, _dp]),
        ("orc_cilia_generalised_forces", None, [v, _dp]),
        ("orc_cilia_segment_velocities", None, [v, _dp]),
        ("orc_cilia_interp_gen_force", C.c_double, [v, C.c_double, C.c_double]),
    ]
    for name, res, args in sig:
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    lib._cilia_ready = True
    return lib


def _opt(a):
    return _p(a) if a is not None else None


class Cilia:
    """Oracle prescribed-cilia stepper for a geometry.CiliaConfig."""

    def __init__(self, cfg, nthreads=None):
        self.lib = _cilia_lib()
        a = cfg.arrays()
        self._keep = a
        ng = 0 if a["gen_force_refs"] is None else a["gen_force_refs"].size
        self.p = OrcCiliaParams(
            cfg.n_swimmers, a["nfil"].ctypes.data_as(C.POINTER(C.c_int)),
            a["nblob"].ctypes.data_as(C.POINTER(C.c_int)), cfg.nseg, int(bool(cfg.interact)),
            _p(a["blob_refs"]), _p(a["fil_refs"]), _p(a["body_x"]), _opt(a["phase"]),
            _opt(a["omega0"]), cfg.r_over_l, cfg.mu, cfg.a_seg, cfg.a_blob, cfg.dt, cfg.beat,
            cfg.amplitude, cfg.recovery_window, cfg.effective_stroke,
            cfg.zero_velocity_avoidance, int(bool(cfg.dynamic_phase)), int(bool(cfg.fixed_body)),
            _opt(a["body_velocity"]), _opt(a["gen_force_refs"]), ng, cfg.gmres_tol,
            cfg.gmres_max_iter, int(bool(cfg.wall)))
        self.h = self.lib.orc_cilia_create(C.byref(self.p), nthreads or nthreads_default())
        if not self.h:
            raise ValueError("orc_cilia_create rejected the configuration")
        self.cfg = cfg
        self.n = self.lib.orc_cilia_system_size(self.h)

    def close(self):
        if self.h:
            self.lib.orc_cilia_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nt):
        it = C.c_int(0)
        rc = self.lib.orc_cilia_step(self.h, nt, C.byref(it))
        return rc, it.value

    def initial_guess(self, nt=0):
        self.lib.orc_cilia_initial_guess(self.h, nt)

    def compute_velocities(self, nt=0):
        it = C.c_int(0)
        rc = self.lib.orc_cilia_compute_velocities(self.h, nt, C.byref(it))
        return rc, it.value

    def integrate_bodies(self):
        self.lib.orc_cilia_integrate_bodies(self.h)

    def end_of_step(self, nt=0):
        self.lib.orc_cilia_end_of_step(self.h, nt)

    def rhs(self):
        out = np.zeros(self.n)
        self.lib.orc_cilia_rhs(self.h, _p(out))
        return out

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.n)
        self.lib.orc_cilia_apply_operator(self.h, _p(x), _p(out))
        return out

    def apply_preconditioner(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.n)
        self.lib.orc_cilia_apply_preconditioner(self.h, _p(x), _p(out))
        return out

    def precon_matrices(self):
        out = np.zeros(36 * self.cfg.n_swimmers)
        self.lib.orc_cilia_precon_matrices(self.h, _p(out))
        return out.reshape(-1, 6, 6)

    def state(self):
        c = self.cfg
        x = np.zeros(3 * c.ns); ph = np.zeros(c.nf); sp = np.zeros(c.nf)
        body = np.zeros(7 * c.n_swimmers); k = np.zeros(3 * c.ns); fq = np.zeros(4 * c.nf)
        self.lib.orc_cilia_get_state(self.h, _p(x), _p(ph), _p(sp), _p(body), _p(k), _p(fq))
        return dict(x=x.reshape(-1, 3), phase=ph, speed=sp, body=body.reshape(-1, 7),
                    vel_dir=k.reshape(-1, 3), fil_q=fq.reshape(-1, 4))

    def solution(self):
        c = self.cfg
        fs = np.zeros(3 * c.ns); fb = np.zeros(3 * c.nb); vb = np.zeros(6 * c.n_swimmers)
        self.lib.orc_cilia_get_solution(self.h, _p(fs), _p(fb), _p(vb))
        return dict(f_segs=fs.reshape(-1, 3), f_blobs=fb.reshape(-1, 3), v_bodies=vb.reshape(-1, 6))

    def generalised_forces(self):
        out = np.zeros(self.cfg.nf)
        self.lib.orc_cilia_generalised_forces(self.h, _p(out))
        return out

    def segment_velocities(self):
        out = np.zeros(3 * self.cfg.ns)
        self.lib.orc_cilia_segment_velocities(self.h, _p(out))
        return out.reshape(-1, 3)

    def interp_gen_force(self, phase, omega0=2.0 * np.pi):
        return self.lib.orc_cilia_interp_gen_force(self.h, float(phase), float(omega0))


def cilia_reference_gen_forces(cfg, steps_per_period, nthreads=None):
    """The reference generalised forces of a beat: one filament on the
    configuration's body, phase 0, phase speed 2 pi, no phase dynamics, for
    one period (config.hpp:259-264, cilia_sim_main.cpp:184-246): vel_dir . f
    after each step's solve."""
    import dataclasses
    one = dataclasses.replace(
        cfg, nfil=np.array([1]), nblob=np.array([cfg.nblob[0]]),
        blob_refs=np.asarray(cfg.blob_refs).reshape(-1, 3)[:cfg.nblob[0]],
        fil_refs=np.asarray(cfg.fil_refs).reshape(-1, 3)[:1], body_x=np.zeros((1, 3)),
        phase=np.zeros(1), omega0=np.array([2.0 * np.pi]), dt=1.0 / steps_per_period,
        dynamic_phase=False, gen_force_refs=None, body_velocity=None)
    o = Cilia(one, nthreads)
    out = np.zeros(steps_per_period)
    for nt in range(steps_per_period):
        o.initial_guess(nt)
        rc, _ = o.compute_velocities(nt)
        assert rc == 0
        out[nt] = o.generalised_forces()[0]
        o.integrate_bodies()
        o.end_of_step(nt)
    return out
