"""ctypes binding of the C ABI in include/gmi.h.

This is the binding a Python caller of the library uses; it carries no
compute of its own.  Loading fails loudly when libgmi.so has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgmi.so")

GMI_OK = 0
GMI_ERR_ARG = -1
GMI_ERR_HIP = -2
GMI_ERR_STATE = -3
GMI_ERR_NOT_FINITE = -4
GMI_ERR_SINGULAR = -5
GMI_ERR_TOO_LARGE = -7
GMI_NOALLOC_MAX_TILES = 4096

GMI_FP64 = 0
GMI_BLOCK_SS, GMI_BLOCK_SB, GMI_BLOCK_BS, GMI_BLOCK_BB = 1, 2, 4, 8
GMI_SEED_UNIFORM, GMI_SEED_EQUATORIAL, GMI_SEED_PLATY = 0, 1, 2
GMI_PRECOND_SELF_MOBILITY, GMI_PRECOND_BODY_INVERSE = 0, 1
WARMSTART_NIN = 51   # GMI_WARMSTART_NIN
GMI_FP32_KAHAN = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class GmiParams(C.Structure):
    _fields_ = [
        ("n_swimmers", C.c_int),
        ("nseg_per_swimmer", _ip),
        ("nblob_per_swimmer", _ip),
        ("nseg_per_filament", C.c_int),
        ("swimmers_interact", C.c_int),
        ("mu", C.c_double),
        ("a_seg", C.c_double),
        ("a_blob", C.c_double),
        ("precision", C.c_int),
        ("device", C.c_int),
    ]


# (name, restype, argtypes) of every exported function declared in gmi.h
SIGNATURES = [
    ("gmi_device_count", C.c_int, [_ip]),
    ("gmi_create", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(GmiParams)]),
    ("gmi_destroy", None, [C.c_void_p]),
    ("gmi_last_error", C.c_char_p, [C.c_void_p]),
    ("gmi_set_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_num_segments", C.c_int, [C.c_void_p]),
    ("gmi_num_blobs", C.c_int, [C.c_void_p]),
    ("gmi_copy_segment_positions_to_device", C.c_int, [C.c_void_p, _dp]),
    ("gmi_copy_segment_forces_to_device", C.c_int, [C.c_void_p, _dp]),
    ("gmi_copy_blob_positions_to_device", C.c_int, [C.c_void_p, _dp]),
    ("gmi_copy_blob_forces_to_device", C.c_int, [C.c_void_p, _dp]),
    ("gmi_copy_interparticle_blob_forces_to_host", C.c_int, [C.c_void_p, _dp]),
    ("gmi_copy_blob_velocities_to_host", C.c_int, [C.c_void_p, _dp]),
    ("gmi_copy_segment_velocities_to_host", C.c_int, [C.c_void_p, _dp]),
    ("gmi_apply_interparticle_forces", C.c_int, [C.c_void_p, C.c_double]),
    ("gmi_wait_for_device", C.c_int, [C.c_void_p]),
    ("gmi_evaluate_segment_segment_mobility", C.c_int, [C.c_void_p]),
    ("gmi_evaluate_segment_blob_mobility", C.c_int, [C.c_void_p]),
    ("gmi_evaluate_blob_blob_mobility", C.c_int, [C.c_void_p]),
    ("gmi_evaluate_blob_segment_mobility", C.c_int, [C.c_void_p]),
    ("gmi_evaluate_grand_mobility", C.c_int, [C.c_void_p]),
    ("gmi_timer_start", C.c_int, [C.c_void_p]),
    ("gmi_timer_stop", C.c_int, [C.c_void_p, _dp]),
    ("gmi_profile_begin", C.c_int, [C.c_void_p]),
    ("gmi_profile_end", C.c_int, [C.c_void_p, _dp, _ip]),
    ("gmi_profile_krylov", C.c_int, [C.c_void_p, _dp, _dp, _ip]),
    ("gmi_set_launch_shape", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("gmi_autotune", C.c_int, [C.c_void_p, C.c_int, _ip, _ip]),
    ("gmi_autotune_blocks", C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, _ip]),
    ("gmi_set_wall", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_evaluate_grand_mobility_noalloc", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
      C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p]),
    # gmi_sim.h
    ("gmi_sim_create", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gmi_make_body_reference_matrices", C.c_int, [C.c_void_p]),
    ("gmi_initial_guess", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_forces_and_torques", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_compute_velocities", C.c_int, [C.c_void_p, C.c_int, _ip]),
    ("gmi_compute_errors", C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, _ip]),
    ("gmi_prepare_jacobian_inv", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_broyden_find_update", C.c_int, [C.c_void_p]),
    ("gmi_apply_update", C.c_int, [C.c_void_p]),
    ("gmi_broyden_end_of_iter", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_step", C.c_int, [C.c_void_p, C.c_int, _ip, _ip]),
    ("gmi_get_state", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    ("gmi_get_mobility", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("gmi_get_broyden", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("gmi_sim_set_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_sim_set_graphs", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("gmi_sim_graph_stats", C.c_int, [C.c_void_p, _ip, _ip]),
    ("gmi_sim_set_gmres_history", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_sim_set_gram_schmidt", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_sim_set_jacobian_pivoting", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_sim_set_body_preconditioner", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_warmstart_init", C.c_int, [C.c_void_p, C.c_int, C.c_ulonglong]),
    ("gmi_warmstart_configure", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    ("gmi_warmstart_dataset_size", C.c_int, [C.c_void_p, _ip]),
    ("gmi_warmstart_dataset", C.c_int, [C.c_void_p, C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.c_int]),
    ("gmi_warmstart_stats", C.c_int, [C.c_void_p, _ip, _ip]),
    ("gmi_seed_sphere", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _ip]),
    ("gmi_sim_write_par", C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int]),
    ("gmi_sim_write_reference_positions", C.c_int, [C.c_void_p, C.c_char_p]),
    ("gmi_sim_write_data", C.c_int, [C.c_void_p, C.c_int, C.c_char_p]),
    ("gmi_sim_write_backup", C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int]),
    ("gmi_sim_read_backup", C.c_int, [C.c_void_p, C.c_char_p, _ip, _ip]),
    ("gmi_warmstart_train", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double,
                                      C.c_ulonglong, _dp]),
    ("gmi_warmstart_param_count", C.c_int, [C.c_void_p, _ip]),
    ("gmi_warmstart_params", C.c_int, [C.c_void_p, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]),
    ("gmi_warmstart_forward", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int,
                                        C.POINTER(C.c_float)]),
    ("gmi_warmstart_loss_grad", C.c_int, [C.c_void_p, C.POINTER(C.c_float),
                                          C.POINTER(C.c_float), C.c_int, _dp,
                                          C.POINTER(C.c_float)]),
    # gmi_cilia.h
    ("gmi_cilia_create", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gmi_cilia_step", C.c_int, [C.c_void_p, C.c_int, _ip]),
    ("gmi_cilia_initial_guess", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_cilia_compute_velocities", C.c_int, [C.c_void_p, C.c_int, _ip]),
    ("gmi_cilia_integrate_bodies", C.c_int, [C.c_void_p]),
    ("gmi_cilia_end_of_step", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_cilia_get_state", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    ("gmi_cilia_get_solution", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("gmi_cilia_generalised_forces", C.c_int, [C.c_void_p, _dp]),
    ("gmi_cilia_segment_velocities", C.c_int, [C.c_void_p, _dp]),
    ("gmi_cilia_system_size", C.c_int, [C.c_void_p, _ip]),
    ("gmi_cilia_rhs", C.c_int, [C.c_void_p, _dp]),
    ("gmi_cilia_apply_operator", C.c_int, [C.c_void_p, _dp, _dp]),
    ("gmi_cilia_apply_preconditioner", C.c_int, [C.c_void_p, _dp, _dp]),
    ("gmi_cilia_set_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("gmi_cilia_set_graphs", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("gmi_cilia_set_gram_schmidt", C.c_int, [C.c_void_p, C.c_int]),
]

GMI_ERR_NOT_CONVERGED = -6

_lib = None


def load(path: str = LIB_PATH):
    """Load libgmi.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing; build it with "
            "`python -m generic_mobility_interface_class_amd.build`")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def device_count() -> int:
    n = C.c_int(0)
    load().gmi_device_count(C.byref(n))
    return int(n.value)


class GmiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gmi error {code}: {msg}")
        self.code = code


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Context:
    """One device context (gmi_create / gmi_destroy)."""

    def __init__(self, nseg_per_swimmer, nblob_per_swimmer, nseg_per_filament,
                 mu=1.0, a_seg=1.0, a_blob=1.0, precision=GMI_FP64,
                 interact=True, device=0):
        self.lib = load()
        ns = np.ascontiguousarray(nseg_per_swimmer, dtype=np.int32)
        nb = np.ascontiguousarray(nblob_per_swimmer, dtype=np.int32)
        self._ns, self._nb = ns, nb
        p = GmiParams(len(ns), ns.ctypes.data_as(_ip), nb.ctypes.data_as(_ip),
                      int(nseg_per_filament), int(bool(interact)), float(mu),
                      float(a_seg), float(a_blob), int(precision), int(device))
        h = C.c_void_p()
        rc = self.lib.gmi_create(C.byref(h), C.byref(p))
        if rc != GMI_OK:
            msg = self.lib.gmi_last_error(h) if h.value else b"bad parameters"
            if h.value:
                self.lib.gmi_destroy(h)
            raise GmiError(rc, msg.decode())
        self.h = h
        self.Ns = int(ns.sum())
        self.Nb = int(nb.sum())

    def last_error(self):
        return self.lib.gmi_last_error(self.h).decode()

    def check(self, rc):
        if rc != GMI_OK:
            raise GmiError(rc, self.lib.gmi_last_error(self.h).decode())

    def call(self, name, *args):
        self.check(getattr(self.lib, name)(self.h, *args))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gmi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # convenience wrappers -------------------------------------------------
    def upload(self, xs=None, fs=None, xb=None, fb=None):
        if xs is not None:
            self._xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1)
            self.call("gmi_copy_segment_positions_to_device", ptr(self._xs))
        if fs is not None:
            self._fs = np.ascontiguousarray(fs, dtype=np.float64).reshape(-1)
            self.call("gmi_copy_segment_forces_to_device", ptr(self._fs))
        if xb is not None:
            self._xb = np.ascontiguousarray(xb, dtype=np.float64).reshape(-1)
            self.call("gmi_copy_blob_positions_to_device", ptr(self._xb))
        if fb is not None:
            self._fb = np.ascontiguousarray(fb, dtype=np.float64).reshape(-1)
            self.call("gmi_copy_blob_forces_to_device", ptr(self._fb))

    def velocities(self):
        vs = np.zeros(6 * self.Ns)
        vb = np.zeros(3 * self.Nb)
        if self.Ns:
            self.call("gmi_copy_segment_velocities_to_host", ptr(vs))
        if self.Nb:
            self.call("gmi_copy_blob_velocities_to_host", ptr(vb))
        self.call("gmi_wait_for_device")
        return vs.reshape(-1, 6), vb.reshape(-1, 3)

    def set_precision(self, p):
        self.call("gmi_set_precision", int(p))

    def autotune(self, reps=3):
        t, ch = C.c_int(0), C.c_int(0)
        self.call("gmi_autotune", int(reps), C.byref(t), C.byref(ch))
        return t.value, ch.value

    def set_wall(self, wall=True):
        self.call("gmi_set_wall", int(bool(wall)))

    def seed_sphere(self, n, kind=0, r_over_l=1.0, max_iterations=200000):
        """Unit-sphere seeding as the reference does (gmi_seed_sphere);
        returns (points [n, 3], iterations)."""
        out = np.zeros(3 * int(n))
        it = C.c_int(0)
        self.call("gmi_seed_sphere", int(n), int(kind), float(r_over_l), int(max_iterations),
                  ptr(out), C.byref(it))
        return out.reshape(-1, 3), it.value

    def autotune_blocks(self, mask, reps=10):
        t, ch = C.c_int(0), C.c_int(0)
        self.call("gmi_autotune_blocks", int(mask), int(reps), C.byref(t), C.byref(ch))
        return t.value, ch.value


class GmiSimParams(C.Structure):
    _fields_ = [
        ("nfil", _ip), ("blob_refs", _dp), ("fil_refs", _dp), ("body_x", _dp),
        ("fil_phase", _dp), ("kb", C.c_double), ("kt", C.c_double), ("dt", C.c_double),
        ("tol", C.c_double), ("jacobian_confidence", C.c_double), ("gmres_tol", C.c_double),
        ("max_broyden_iter", C.c_int), ("gmres_max_iter", C.c_int),
        ("num_euler_steps", C.c_int), ("barrier_K", C.c_double),
        ("curvature_amp", C.c_double), ("curvature_period", C.c_double),
        ("sphere_radius", C.c_double), ("sphere_radii", _dp),
        ("end_force", C.c_double), ("perturbation", _dp),
    ]


class Sim(Context):
    """Device-resident time stepper for a geometry.SimConfig (gmi_sim.h)."""

    def __init__(self, cfg, precision=GMI_FP64, device=0):
        a = cfg.arrays()
        g0 = cfg.swimmers[0]
        super().__init__(a["nfil"] * cfg.nseg, a["nblob"], cfg.nseg, mu=g0.mu,
                         a_seg=g0.a_seg, a_blob=g0.a_blob, precision=precision,
                         interact=bool(cfg.interact), device=device)
        self._a = a
        self.sp = GmiSimParams(
            a["nfil"].ctypes.data_as(_ip), ptr(a["blob_refs"]), ptr(a["fil_refs"]),
            ptr(a["body_x"]), ptr(a["phases"]), cfg.kb, cfg.kt, cfg.dt, cfg.tol,
            cfg.jacobian_confidence, cfg.gmres_tol, cfg.max_broyden_iter, cfg.gmres_max_iter,
            cfg.num_euler_steps, cfg.barrier_K, cfg.curvature_amp, cfg.curvature_period,
            float(a["radius"][0]), ptr(a["radius"]), float(cfg.end_force),
            ptr(a["perturbation"]) if a["perturbation"] is not None else None)
        self.call("gmi_sim_create", C.byref(self.sp))
        self.nsw = len(a["nfil"])
        self.nfil_total = int(a["nfil"].sum())
        self.nseg = cfg.nseg
        self.nbroy = 6 * (self.nfil_total * cfg.nseg + self.nsw)

    def make_body_reference_matrices(self):
        self.call("gmi_make_body_reference_matrices")

    def step(self, nt):
        bi, gi = C.c_int(0), C.c_int(0)
        rc = self.lib.gmi_step(self.h, nt, C.byref(bi), C.byref(gi))
        return rc, bi.value, gi.value

    def state(self):
        S = self.nseg
        x = np.zeros(3 * self.Ns); q = np.zeros(4 * self.Ns); u = np.zeros(3 * self.Ns)
        lam = np.zeros(self.nfil_total * (3 * S + 3)); body = np.zeros(10 * self.nsw)
        self.call("gmi_get_state", ptr(x), ptr(q), ptr(u), ptr(lam), ptr(body))
        return dict(x=x.reshape(-1, 3), q=q.reshape(-1, 4), u=u.reshape(-1, 3),
                    lam=lam.reshape(self.nfil_total, -1), body=body.reshape(-1, 10))

    def mobility(self):
        vs = np.zeros(6 * self.Ns); vb = np.zeros(6 * self.nsw); fb = np.zeros(3 * self.Nb)
        ref = np.zeros(72 * self.nsw)
        self.call("gmi_get_mobility", ptr(vs), ptr(vb), ptr(fb), ptr(ref))
        ref = ref.reshape(self.nsw, 2, 6, 6)
        return dict(v_segs=vs.reshape(-1, 6), v_bodies=vb.reshape(-1, 6),
                    f_blobs=fb.reshape(-1, 3), Nref=ref[:, 0], KTKinv=ref[:, 1])

    def broyden(self, which=0):
        e = np.zeros(self.nbroy)
        self.call("gmi_get_broyden", int(which), ptr(e))
        return e

    def set_sim_precision(self, p):
        self.call("gmi_sim_set_precision", int(p))

    def set_graphs(self, use_graphs=True, chunk=6):
        self.call("gmi_sim_set_graphs", int(bool(use_graphs)), int(chunk))

    def graph_stats(self):
        """(Broyden iterations run from captured graphs, graphs held)."""
        it, ng = C.c_int(0), C.c_int(0)
        self.call("gmi_sim_graph_stats", C.byref(it), C.byref(ng))
        return it.value, ng.value

    def set_body_preconditioner(self, kind=GMI_PRECOND_BODY_INVERSE):
        """Rigid-body GMRES preconditioner: the reference's self-mobility
        diagonal, or the exact body-frame block inverse (kept in HBM)."""
        self.call("gmi_sim_set_body_preconditioner", int(kind))

    def set_gmres_history(self, m=16):
        """GMRES initial guess fitted on the last m right-hand sides (0: the
        reference's previous solution)."""
        self.call("gmi_sim_set_gmres_history", int(m))

    def set_gram_schmidt(self, passes=1):
        """Gram-Schmidt passes of the reference-preconditioned Arnoldi step
        (1: one pass, as the reference; 2: with reorthogonalisation)."""
        self.call("gmi_sim_set_gram_schmidt", int(passes))

    def set_jacobian_pivoting(self, always_pivot=True):
        self.call("gmi_sim_set_jacobian_pivoting", int(bool(always_pivot)))

    # ---- output and checkpoint files (reference formats) --------------
    def write_par(self, name, total_time_steps, plot_frequency):
        self.call("gmi_sim_write_par", str(name).encode(), int(total_time_steps),
                  int(plot_frequency))

    def write_reference_positions(self, name):
        self.call("gmi_sim_write_reference_positions", str(name).encode())

    def write_data(self, nt, name):
        self.call("gmi_sim_write_data", int(nt), str(name).encode())

    def write_backup(self, nt, name, extended=True):
        self.call("gmi_sim_write_backup", int(nt), str(name).encode(), int(bool(extended)))

    def read_backup(self, name):
        nt, exact = C.c_int(0), C.c_int(0)
        self.call("gmi_sim_read_backup", str(name).encode(), C.byref(nt), C.byref(exact))
        return nt.value, bool(exact.value)

    # ---- learned warm start ------------------------------------------
    def warmstart_init(self, hidden=128, seed=0):
        self.call("gmi_warmstart_init", int(hidden), int(seed))
        n = C.c_int(0)
        self.call("gmi_warmstart_param_count", C.byref(n))
        self.nparams = n.value

    def warmstart_configure(self, apply=False, collect=0, stride=1):
        self.call("gmi_warmstart_configure", int(bool(apply)), int(collect), int(stride))

    def warmstart_dataset_size(self):
        n = C.c_int(0)
        self.call("gmi_warmstart_dataset_size", C.byref(n))
        return n.value

    def warmstart_dataset(self, rows=None):
        n = self.warmstart_dataset_size() if rows is None else int(rows)
        X = np.zeros((n, WARMSTART_NIN), np.float32)
        T = np.zeros((n, 6), np.float32)
        fp = C.POINTER(C.c_float)
        self.call("gmi_warmstart_dataset", X.ctypes.data_as(fp), T.ctypes.data_as(fp), n)
        return X, T

    def warmstart_stats(self):
        a, f = C.c_int(0), C.c_int(0)
        self.call("gmi_warmstart_stats", C.byref(a), C.byref(f))
        return a.value, f.value

    def warmstart_train(self, iters=2000, batch=1024, lr=1e-3, seed=0):
        loss = C.c_double(0.0)
        self.call("gmi_warmstart_train", int(iters), int(batch), float(lr), int(seed),
                  C.byref(loss))
        return loss.value

    def warmstart_params(self, set_to=None):
        out = np.zeros(self.nparams, np.float32)
        src = None
        if set_to is not None:
            src = np.ascontiguousarray(set_to, np.float32)
        self.call("gmi_warmstart_params", out.ctypes.data_as(C.POINTER(C.c_float)),
                  None if src is None else src.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def warmstart_forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 6), np.float32)
        fp = C.POINTER(C.c_float)
        self.call("gmi_warmstart_forward", X.ctypes.data_as(fp), X.shape[0], Y.ctypes.data_as(fp))
        return Y

    def warmstart_loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_double(0.0)
        fp = C.POINTER(C.c_float)
        self.call("gmi_warmstart_loss_grad", X.ctypes.data_as(fp), T.ctypes.data_as(fp),
                  X.shape[0], C.byref(loss), g.ctypes.data_as(fp))
        return loss.value, g


# ---- prescribed cilia (gmi_cilia.h) ------------------------------------------

GMI_BEAT_BUILD, GMI_BEAT_FULFORD_BLAKE, GMI_BEAT_CORAL_LARVAE = 0, 1, 2


class GmiCiliaParams(C.Structure):
    _fields_ = [
        ("nfil", _ip), ("blob_refs", _dp), ("fil_refs", _dp), ("body_x", _dp), ("phase", _dp),
        ("omega0", _dp), ("r_over_l", C.c_double), ("dt", C.c_double), ("beat", C.c_int),
        ("amplitude", C.c_double), ("recovery_window", C.c_double),
        ("effective_stroke", C.c_double), ("zero_velocity_avoidance", C.c_double),
        ("dynamic_phase", C.c_int), ("fixed_body", C.c_int), ("body_velocity", _dp),
        ("gen_force_refs", _dp), ("n_gen_force_refs", C.c_int), ("gmres_tol", C.c_double),
        ("gmres_max_iter", C.c_int),
    ]


def _optp(a):
    return ptr(a) if a is not None else None


class Cilia(Context):
    """Prescribed-cilia time stepper for a geometry.CiliaConfig (gmi_cilia.h)."""

    def __init__(self, cfg, precision=GMI_FP64, device=0):
        a = cfg.arrays()
        super().__init__(a["nfil"] * cfg.nseg, a["nblob"], cfg.nseg, mu=cfg.mu,
                         a_seg=cfg.a_seg, a_blob=cfg.a_blob, precision=precision,
                         interact=cfg.interact, device=device)
        self._a = a
        ng = 0 if a["gen_force_refs"] is None else a["gen_force_refs"].size
        self.cp = GmiCiliaParams(
            a["nfil"].ctypes.data_as(_ip), _optp(a["blob_refs"]), ptr(a["fil_refs"]),
            _optp(a["body_x"]), _optp(a["phase"]), _optp(a["omega0"]), cfg.r_over_l, cfg.dt,
            cfg.beat, cfg.amplitude, cfg.recovery_window, cfg.effective_stroke,
            cfg.zero_velocity_avoidance, int(bool(cfg.dynamic_phase)), int(bool(cfg.fixed_body)),
            _optp(a["body_velocity"]), _optp(a["gen_force_refs"]), ng, cfg.gmres_tol,
            cfg.gmres_max_iter)
        if getattr(cfg, "wall", False):
            self.set_wall(True)
        self.call("gmi_cilia_create", C.byref(self.cp))
        self.cfg = cfg
        n = C.c_int(0)
        self.call("gmi_cilia_system_size", C.byref(n))
        self.n = n.value

    def step(self, nt):
        it = C.c_int(0)
        rc = self.lib.gmi_cilia_step(self.h, int(nt), C.byref(it))
        return rc, it.value

    def initial_guess(self, nt=0):
        self.call("gmi_cilia_initial_guess", int(nt))

    def compute_velocities(self, nt=0):
        it = C.c_int(0)
        rc = self.lib.gmi_cilia_compute_velocities(self.h, int(nt), C.byref(it))
        return rc, it.value

    def integrate_bodies(self):
        self.call("gmi_cilia_integrate_bodies")

    def end_of_step(self, nt=0):
        self.call("gmi_cilia_end_of_step", int(nt))

    def state(self):
        c = self.cfg
        x = np.zeros(3 * c.ns); ph = np.zeros(c.nf); sp = np.zeros(c.nf)
        body = np.zeros(7 * c.n_swimmers); k = np.zeros(3 * c.ns)
        self.call("gmi_cilia_get_state", ptr(x), ptr(ph), ptr(sp), ptr(body), ptr(k))
        return dict(x=x.reshape(-1, 3), phase=ph, speed=sp, body=body.reshape(-1, 7),
                    vel_dir=k.reshape(-1, 3))

    def solution(self):
        c = self.cfg
        fs = np.zeros(3 * c.ns); fb = np.zeros(3 * c.nb); vb = np.zeros(6 * c.n_swimmers)
        self.call("gmi_cilia_get_solution", ptr(fs), ptr(fb) if c.nb else None, ptr(vb))
        return dict(f_segs=fs.reshape(-1, 3), f_blobs=fb.reshape(-1, 3), v_bodies=vb.reshape(-1, 6))

    def generalised_forces(self):
        out = np.zeros(self.cfg.nf)
        self.call("gmi_cilia_generalised_forces", ptr(out))
        return out

    def segment_velocities(self):
        out = np.zeros(3 * self.cfg.ns)
        self.call("gmi_cilia_segment_velocities", ptr(out))
        return out.reshape(-1, 3)

    def rhs(self):
        out = np.zeros(self.n)
        self.call("gmi_cilia_rhs", ptr(out))
        return out

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.n)
        self.call("gmi_cilia_apply_operator", ptr(x), ptr(out))
        return out

    def apply_preconditioner(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.n)
        self.call("gmi_cilia_apply_preconditioner", ptr(x), ptr(out))
        return out

    def set_cilia_precision(self, p):
        self.call("gmi_cilia_set_precision", int(p))

    def set_graphs(self, use_graphs=True, chunk=6):
        self.call("gmi_cilia_set_graphs", int(bool(use_graphs)), int(chunk))

    def set_gram_schmidt(self, passes=1):
        """Gram-Schmidt passes per Arnoldi step (1: one, as the reference; 2:
        reorthogonalised)."""
        self.call("gmi_cilia_set_gram_schmidt", int(passes))
