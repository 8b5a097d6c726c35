"""Deterministic synthetic swimmer geometry and inputs.

The reference seeds its sphere surfaces by a spiral start plus repulsive
relaxation (seeding.cu:13-389) and draws filament perturbations from
std::random_device (filament.cpp:345-360).  Neither is reproducible, so the
workloads here use closed-form geometry instead:

* blob centres: vertices of a geodesic icosphere of frequency f, which has
  10 f^2 + 2 vertices (642 -> f=8, 2562 -> 16, 4002 -> 20, 10242 -> 32,
  16002 -> 40);
* filament bases: Fibonacci-sphere points at radius
  R + 1.05 (DL/2 + a_blob), the offset the reference uses between the blob
  shell and the filament base sphere (swimmer.cpp:171 vs rigid_body.cpp:186);
* filaments: straight, normal to the surface, segment k at base + k DL n
  (filament.cpp:297-307);
* random numbers: splitmix64-seeded xoshiro256**, normals by Box-Muller,
  so every platform produces the same bits.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

_M64 = (1 << 64) - 1


def _splitmix64(state: int):
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


class Xoshiro256ss:
    """xoshiro256** (Blackman & Vigna 2018), seeded through splitmix64."""

    def __init__(self, seed: int):
        st = seed & _M64
        s = []
        for _ in range(4):
            st, z = _splitmix64(st)
            s.append(z)
        self.s = s

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & _M64

    def next_u64(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1) from the top 53 bits."""
        out = np.empty(n, dtype=np.float64)
        for i in range(n):
            out[i] = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return out

    def normal(self, n: int) -> np.ndarray:
        """n N(0,1) doubles by Box-Muller (pairs: cos branch, then sin)."""
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        u1 = 1.0 - u[0::2]          # (0, 1]
        u2 = u[1::2]
        rad = np.sqrt(-2.0 * np.log(u1))
        z = np.empty(2 * m)
        z[0::2] = rad * np.cos(2.0 * math.pi * u2)
        z[1::2] = rad * np.sin(2.0 * math.pi * u2)
        return z[:n]


def icosahedron():
    p = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
                  [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
                  [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], float)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    return v, f


def geodesic_sphere(freq: int, radius: float = 1.0) -> np.ndarray:
    """Vertices of the frequency-`freq` geodesic icosphere (10 f^2 + 2 points).

    Each icosahedron face is split into f^2 triangles on a barycentric grid;
    the grid points are projected to the sphere.  Shared edge and corner
    points are emitted once, in first-visit order, so the order is fixed.
    """
    if freq < 1:
        raise ValueError("freq must be >= 1")
    v, faces = icosahedron()
    pts = []
    seen = {}
    for (a, b, c) in faces:
        A, B, C = v[a], v[b], v[c]
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                k = freq - i - j
                # exact integer key of the barycentric point on the shared
                # edge/corner: sort contributions by vertex id
                key = tuple(sorted(((a, i), (b, j), (c, k)), key=lambda t: t[0]))
                key = tuple(t for t in key if t[1] != 0)
                if key in seen:
                    continue
                p = (i * A + j * B + k * C) / freq
                p = p / np.linalg.norm(p)
                seen[key] = len(pts)
                pts.append(p)
    out = np.array(pts) * radius
    assert out.shape[0] == 10 * freq * freq + 2
    return out


def geodesic_freq_for(nblob: int) -> int:
    f = int(round(math.sqrt((nblob - 2) / 10.0)))
    if 10 * f * f + 2 != nblob:
        raise ValueError(f"{nblob} is not a geodesic count 10 f^2 + 2")
    return f


def fibonacci_sphere(n: int) -> np.ndarray:
    """n near-uniform unit vectors (golden-angle spiral)."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)


@dataclass
class SwimmerGeometry:
    nfil: int
    nseg: int
    nblob: int
    R: float                 # blob-shell radius
    a_seg: float = 1.0
    a_blob: float = 1.0
    mu: float = 1.0
    centre: np.ndarray = field(default_factory=lambda: np.zeros(3))
    blob_unit: np.ndarray = None   # unit-sphere blob directions (default: geodesic)
    fil_unit: np.ndarray = None    # unit filament base directions (default: Fibonacci)

    @property
    def DL(self) -> float:   # config.hpp:204
        return 2.2 * self.a_seg

    @property
    def ns(self) -> int:
        return self.nfil * self.nseg

    def blob_references(self) -> np.ndarray:
        if self.blob_unit is not None:
            return np.asarray(self.blob_unit, dtype=np.float64).reshape(-1, 3) * self.R
        return geodesic_sphere(geodesic_freq_for(self.nblob), self.R)

    def filament_directions(self) -> np.ndarray:
        if self.fil_unit is not None:
            u = np.asarray(self.fil_unit, dtype=np.float64).reshape(-1, 3)
            return u / np.linalg.norm(u, axis=1, keepdims=True)
        return fibonacci_sphere(self.nfil)

    def base_radius(self) -> float:
        return self.R + 1.05 * (0.5 * self.DL + self.a_blob)

    def positions(self):
        """(x_segs [Ns,3], x_blobs [Nb,3]) of the straight reference state."""
        dirs = self.filament_directions()
        base = self.centre + dirs * self.base_radius()
        k = np.arange(self.nseg, dtype=np.float64)
        xs = base[:, None, :] + (k[None, :, None] * self.DL) * dirs[:, None, :]
        xb = self.centre + self.blob_references()
        return xs.reshape(-1, 3).copy(), xb


def random_forces(ns: int, nb: int, seed: int, blob_torques: bool = False):
    """F_s [Ns,6] and F_b [Nb,3] (or [Nb,6]) i.i.d. N(0,1).

    Draw order: all segment (f, tau) rows, then all blob rows.
    """
    rng = Xoshiro256ss(seed)
    fs = rng.normal(6 * ns).reshape(ns, 6)
    fb = rng.normal((6 if blob_torques else 3) * nb).reshape(nb, -1)
    return fs, fb


# Workload definitions (sizes of the synthetic configurations)
WORKLOADS = {
    # one swimmer, 962 particles, CPU-runnable apply
    "apply_962": dict(nfil=16, nseg=20, nblob=642, R=10.0),
    # implicit-timestep swimmer: 5120 segments, 4002 blobs
    "step_9122": dict(nfil=256, nseg=20, nblob=4002, R=25.0),
    # large single apply: 30720 segments + 16002 blobs = 46722 particles
    "apply_46722": dict(nfil=1024, nseg=30, nblob=16002, R=40.0),
}


def workload_geometry(name: str) -> SwimmerGeometry:
    return SwimmerGeometry(**WORKLOADS[name])


@dataclass
class SimConfig:
    """Inputs of an implicit-timestep run (one or more swimmers, coupled or
    independent).

    Physics follows the reference defaults (config.hpp:100-154) where the
    workload does not fix them: KB = KT = 1800, Broyden TOL 1e-4 with
    JACOBIAN_CONFIDENCE_FACTOR 0.1 and at most 100 iterations, GMRES at most
    150 iterations, one backward-Euler step before BDF2.  The driving is a
    preferred curvature amp*sin(2 pi t/period + phase_f) with amp = 0.5/L,
    period 1 and dt = period/300; the barrier constant is 0.1*KB (the
    reference's value when no end force defines it, config.hpp:244-247).
    """
    swimmers: list            # list of SwimmerGeometry
    phases: np.ndarray        # one per filament, all swimmers
    dt: float = 1.0 / 300.0
    kb: float = 1800.0
    kt: float = 1800.0
    tol: float = 1e-4
    jacobian_confidence: float = 0.1
    gmres_tol: float = 1e-8
    max_broyden_iter: int = 100
    gmres_max_iter: int = 150
    num_euler_steps: int = 1
    barrier_factor: float = 0.1       # K = barrier_factor * kb
    curvature_period: float = 1.0
    curvature_amp: float = None       # default 0.5 / L
    # the reference's default CILIA_TYPE 0 physics: a follower force of this
    # magnitude at every filament tip (END_FORCE_MAGNITUDE) and, at nt = 0,
    # per filament the two perturbation forces on segments 4 and 12
    # ((n_filaments, 6) array, or None)
    end_force: float = 0.0
    perturbation: np.ndarray = None
    # swimmers couple hydrodynamically (every particle with every other, the
    # reference's NSWIM > 1 run) or are independent (a batch)
    interact: bool = False

    def __post_init__(self):
        g0 = self.swimmers[0]
        if self.curvature_amp is None:
            self.curvature_amp = 0.5 / (g0.nseg * g0.DL)
        for g in self.swimmers:
            if (g.nseg, g.a_seg, g.a_blob, g.mu) != (g0.nseg, g0.a_seg, g0.a_blob, g0.mu):
                raise ValueError("swimmers must share NSEG, radii and viscosity")

    @property
    def nseg(self):
        return self.swimmers[0].nseg

    @property
    def barrier_K(self):
        return self.barrier_factor * self.kb

    def arrays(self):
        """Flat arrays shared by the library and the oracle."""
        nfil = np.array([g.nfil for g in self.swimmers], np.int32)
        nblob = np.array([g.nblob for g in self.swimmers], np.int32)
        blob_refs = np.concatenate([g.blob_references().reshape(-1) for g in self.swimmers])
        fil_refs = np.concatenate([(g.filament_directions() * g.base_radius()).reshape(-1)
                                   for g in self.swimmers])
        body_x = np.concatenate([np.asarray(g.centre, float) for g in self.swimmers])
        radius = np.array([g.R for g in self.swimmers], dtype=np.float64)
        pert = None
        if self.perturbation is not None:
            pert = np.ascontiguousarray(self.perturbation, dtype=np.float64).reshape(-1)
            assert pert.size == 6 * int(nfil.sum())
        return dict(nfil=nfil, nblob=nblob, blob_refs=np.ascontiguousarray(blob_refs),
                    fil_refs=np.ascontiguousarray(fil_refs), body_x=body_x,
                    phases=np.ascontiguousarray(self.phases, dtype=np.float64), radius=radius,
                    perturbation=pert)


def sim_config(nfil, nseg, nblob, R, seed=0, **kw):
    """Single swimmer with random driving phases U[0, 2 pi) from `seed`."""
    g = SwimmerGeometry(nfil=nfil, nseg=nseg, nblob=nblob, R=R)
    rng = Xoshiro256ss(seed)
    phases = 2.0 * math.pi * rng.uniform(nfil)
    return SimConfig(swimmers=[g], phases=phases, **kw)


SIM_WORKLOADS = {
    # 16 x 20 segments + 642 blobs (the 962-particle swimmer)
    "step_962": dict(nfil=16, nseg=20, nblob=642, R=10.0),
    # 256 x 20 segments + 4002 blobs
    "step_9122": dict(nfil=256, nseg=20, nblob=4002, R=25.0),
}


def sim_workload(name, seed=0, **kw):
    return sim_config(seed=seed, **SIM_WORKLOADS[name], **kw)


def end_force_magnitude(nseg, kb=1800.0, a_seg=1.0, dimensionless_force=220.0):
    """END_FORCE_MAGNITUDE = DIMENSIONLESS_FORCE KB / (DL^2 NSEG^2)
    (config.hpp:108-110, 239)."""
    DL = 2.2 * a_seg
    return dimensionless_force * kb / (DL * DL * nseg * nseg)


def perturbations(nfil, magnitude, seed):
    """Per filament the two initial perturbation forces of the reference
    (filament.cpp:336-360): magnitude 0.01 END_FORCE_MAGNITUDE along
    directions of normalised N(0,1) triples.  The reference draws them from
    std::random_device; here they come from the project RNG (draw order:
    filament by filament, perturbation1 then perturbation2)."""
    rng = Xoshiro256ss(seed)
    out = np.zeros((nfil, 6))
    for f in range(nfil):
        for k in range(2):
            d = rng.normal(3)
            out[f, 3 * k:3 * k + 3] = magnitude * d / np.linalg.norm(d)
    return out


def instability_config(swimmers, seed=0, **kw):
    """The reference's default physics (CILIA_TYPE 0, config.hpp:13,108-110,
    239-241): follower end force END_FORCE_MAGNITUDE at every filament tip,
    barrier K = REPULSIVE_FORCE_FACTOR (2) x END_FORCE_MAGNITUDE, DT =
    36.3833/STEPS_PER_PERIOD with 300 steps per period, no preferred
    curvature, GMRES tolerance LINEAR_SYSTEM_TOL = 1e-10, and the nt = 0
    perturbations (0.01 END_FORCE_MAGNITUDE) from `seed`."""
    g0 = swimmers[0]
    kb = kw.pop("kb", 1800.0)
    ef = end_force_magnitude(g0.nseg, kb, g0.a_seg)
    nf = sum(g.nfil for g in swimmers)
    args = dict(dt=36.3833 / 300.0, kb=kb, kt=kb, curvature_amp=0.0, gmres_tol=1e-10,
                barrier_factor=2.0 * ef / kb, end_force=ef,
                perturbation=perturbations(nf, 0.01 * ef, seed))
    args.update(kw)
    return SimConfig(swimmers=list(swimmers), phases=np.zeros(nf), **args)


def ensemble_config(n_swimmers=64, seed=0, nfil_min=16, nfil_max=512, nseg=20,
                    freq_min=8, freq_max=32, spacing=None, **kw):
    """Batch of hydrodynamically independent swimmers of uneven size.

    Per swimmer (draw order: all NFIL, then all phases):
      * NFIL = round(exp(U[ln nfil_min, ln nfil_max])), log-uniform;
      * blobs proportional to NFIL at the density of the 256-filament,
        4002-blob swimmer (4002 * NFIL / 256), snapped to the nearest geodesic
        count 10 f^2 + 2 with f in [freq_min, freq_max] (642 .. 10242);
      * blob-shell radius R = 25 sqrt(Nb / 4002): the blob areal density of
        the 4002-blob swimmer on R = 25 (and of 642 blobs on R = 10);
      * centres on the x axis, `spacing` apart (default: 2 x the largest
        swimmer's reach), so the swimmers never touch;
      * driving phases U[0, 2 pi) per filament.
    """
    rng = Xoshiro256ss(seed)
    lo, hi = math.log(nfil_min), math.log(nfil_max)
    nfil = [int(round(math.exp(lo + (hi - lo) * u))) for u in rng.uniform(n_swimmers)]
    geoms = []
    for nf in nfil:
        f = int(round(math.sqrt(max(4002.0 * nf / 256.0 - 2.0, 0.0) / 10.0)))
        f = min(max(f, freq_min), freq_max)
        nb = 10 * f * f + 2
        geoms.append(SwimmerGeometry(nfil=nf, nseg=nseg, nblob=nb, R=25.0 * math.sqrt(nb / 4002.0)))
    reach = max(g.base_radius() + nseg * g.DL for g in geoms)
    step = spacing if spacing is not None else 2.0 * reach + 10.0
    for k, g in enumerate(geoms):
        g.centre = np.array([k * step, 0.0, 0.0])
    phases = 2.0 * math.pi * rng.uniform(sum(nfil))
    return SimConfig(swimmers=geoms, phases=phases, **kw)


ENSEMBLE_WORKLOADS = {
    # 64 swimmers, NFIL log-uniform in [16, 512], 642..10242 blobs
    "ensemble_64": dict(n_swimmers=64, nfil_min=16, nfil_max=512),
    # 6 swimmers, NFIL in [2, 12]: small enough for the oracle
    "ensemble_6": dict(n_swimmers=6, nfil_min=2, nfil_max=12, nseg=10, freq_min=3, freq_max=8),
}


def ensemble_workload(name, seed=0, **kw):
    return ensemble_config(seed=seed, **ENSEMBLE_WORKLOADS[name], **kw)


# ---------------------------------------------------------------- reference seeding

SEED_KINDS = {"uniform": 0, "equatorial": 1, "platy": 2}


def seed_file_name(n, kind):
    """The reference's cache file of a seeding (seeding.cu:121,352-356)."""
    return {"uniform": f"sphere{n}.seed", "equatorial": f"sphere{n}_equatorial.seed",
            "platy": f"sphere{n}_platy.seed"}[kind]


def read_seed_file(path, n):
    """Unit-sphere points from a seed file (rigid_body.cpp:214-218,
    swimmer.cpp:188-194: 3N whitespace-separated numbers)."""
    vals = np.array([float(t) for t in open(path).read().split()])
    if vals.size < 3 * n:
        raise ValueError(f"{path}: {vals.size} values, need {3 * n}")
    return vals[:3 * n].reshape(n, 3)


def write_seed_file(path, points, per_line):
    """As the reference writes it: "%g " after every value (blob seeds,
    seeding.cu:121-126) or "%g \n" (filament seeds, seeding.cu:361-365)."""
    sep = " \n" if per_line else " "
    with open(path, "w") as f:
        f.write("".join(("%g" % v) + sep for v in np.ravel(points)))


def reference_seed(ctx, n, kind="uniform", r_over_l=1.0, cache_dir=None, per_line=False):
    """Unit-sphere points of the reference's seeding.  With cache_dir, an
    existing sphere<N>*.seed file there is read (6 significant digits, as
    the reference reads its cache); otherwise the library seeds on the GPU
    (gmi_seed_sphere via `ctx`, a capi.Context) and, with cache_dir, writes
    the file (the platy kind has no cache file in the reference)."""
    path = os.path.join(cache_dir, seed_file_name(n, kind)) if cache_dir else None
    if path and kind != "platy" and os.path.exists(path):
        return read_seed_file(path, n)
    if ctx is None:
        raise RuntimeError("no seed file and no device context to compute it")
    pts, _ = ctx.seed_sphere(n, SEED_KINDS[kind], r_over_l)
    if path and kind != "platy":
        write_seed_file(path, pts, per_line)
    return pts


def reference_swimmer(ctx, nfil, nseg, nblob, r_over_l=1.0, a_seg=1.0, a_blob=1.4,
                      seeding="uniform", cache_dir=None, centre=None):
    """A swimmer laid out as the reference lays out a spherical body: blobs
    on a sphere of radius R_OVER_L * L from the uniform seeding
    (rigid_body.cpp:186-226), filament bases from the chosen seeding at
    R_OVER_L * L + 1.05 (DL/2 + RBLOB) (swimmer.cpp:171-206).  Defaults are
    the reference's RSEG = 1, RBLOB = 1.4 (config.hpp:102-103)."""
    L = nseg * 2.2 * a_seg
    blobs = reference_seed(ctx, nblob, "uniform", cache_dir=cache_dir, per_line=False)
    fils = reference_seed(ctx, nfil, seeding, r_over_l, cache_dir=cache_dir, per_line=True)
    return SwimmerGeometry(nfil=nfil, nseg=nseg, nblob=nblob, R=r_over_l * L, a_seg=a_seg,
                           a_blob=a_blob, centre=np.zeros(3) if centre is None else centre,
                           blob_unit=blobs, fil_unit=fils)


# ---------------------------------------------------------------------------
# Prescribed cilia (the reference's CILIA_TYPE 3 build)

BEAT_BUILD, BEAT_FULFORD_BLAKE, BEAT_CORAL_LARVAE = 0, 1, 2


@dataclass
class CiliaConfig:
    """Inputs of a prescribed-cilia run (include/gmi_cilia.h).

    Defaults are the reference's (config.hpp:21-47, 76, 150-154): coral-larva
    beat, dynamic phases, free bodies, DT = 1/STEPS_PER_PERIOD with 300
    steps per period, GMRES tolerance 1e-10 within 150 iterations; the
    build-a-beat parameters are its SCALED_BEAT_AMPLITUDE 2.0, windows 0.25
    and 0.25, zero-velocity avoidance 0.05.
    """
    nseg: int
    nfil: np.ndarray          # per swimmer
    nblob: np.ndarray         # per swimmer
    blob_refs: np.ndarray     # body-frame blob positions, all swimmers [Nb, 3]
    fil_refs: np.ndarray      # body-frame filament bases, all swimmers [Nf, 3]
    body_x: np.ndarray        # initial centres [Nsw, 3]
    phase: np.ndarray = None  # initial phases [Nf] (None: metachronal start)
    omega0: np.ndarray = None # natural frequencies [Nf] (None: 2 pi)
    r_over_l: float = 1.0
    mu: float = 1.0
    a_seg: float = 1.0
    a_blob: float = 1.4
    dt: float = 1.0 / 300.0
    beat: int = BEAT_CORAL_LARVAE
    amplitude: float = 2.0
    recovery_window: float = 0.25
    effective_stroke: float = 0.25
    zero_velocity_avoidance: float = 0.05
    dynamic_phase: bool = True
    fixed_body: bool = False
    body_velocity: np.ndarray = None   # [Nsw, 6] for fixed bodies (None: zero)
    gen_force_refs: np.ndarray = None
    gmres_tol: float = 1e-10
    gmres_max_iter: int = 150
    interact: bool = True
    wall: bool = False        # carpet above a no-slip wall at z = 0

    @property
    def n_swimmers(self):
        return len(self.nfil)

    @property
    def nf(self):
        return int(np.sum(self.nfil))

    @property
    def ns(self):
        return self.nf * self.nseg

    @property
    def nb(self):
        return int(np.sum(self.nblob))

    def system_size(self):
        return (3 * self.ns + 3 * self.nb + (0 if self.fixed_body else 6 * self.n_swimmers)
                + (self.nf if self.dynamic_phase else 0))

    def arrays(self):
        """Flat contiguous arrays shared by the library and the oracle."""
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        return dict(nfil=np.ascontiguousarray(self.nfil, dtype=np.int32),
                    nblob=np.ascontiguousarray(self.nblob, dtype=np.int32),
                    blob_refs=f(self.blob_refs), fil_refs=f(self.fil_refs),
                    body_x=f(self.body_x), phase=f(self.phase), omega0=f(self.omega0),
                    body_velocity=f(self.body_velocity), gen_force_refs=f(self.gen_force_refs))


def cilia_config(nfil, nseg, nblob, n_swimmers=1, r_over_l=1.0, a_seg=1.0, a_blob=1.4,
                 **kw):
    """Spherical swimmers laid out as the reference lays out a prescribed-
    cilia swimmer: blobs on a sphere of radius R_OVER_L DL NSEG
    (rigid_body.cpp:184), filament bases at R_OVER_L L + 1.05 (DL/2 + RBLOB)
    with L = DL (NSEG - 1) (swimmer.cpp:54-60,171), bodies spaced
    3 (R + DL NSEG) apart along x (rigid_body.cpp:188-192).  Blob directions
    are a geodesic sphere and filament directions a Fibonacci sphere (the
    reference's repulsive seeding is gmi_seed_sphere)."""
    DL = 2.2 * a_seg
    Rb = r_over_l * DL * nseg
    Rf = r_over_l * DL * (nseg - 1) + 1.05 * (0.5 * DL + a_blob)
    blobs = geodesic_sphere(geodesic_freq_for(nblob), Rb)
    if blobs.shape[0] != nblob:
        raise ValueError(f"nblob must be a geodesic count 10 f^2 + 2, got {nblob}")
    fils = fibonacci_sphere(nfil) * Rf
    spacing = 3.0 * (Rb + DL * nseg)
    body_x = np.array([[n * spacing, 0.0, 0.0] for n in range(n_swimmers)])
    return CiliaConfig(nseg=nseg, nfil=np.full(n_swimmers, nfil), nblob=np.full(n_swimmers, nblob),
                       blob_refs=np.tile(blobs, (n_swimmers, 1)),
                       fil_refs=np.tile(fils, (n_swimmers, 1)), body_x=body_x,
                       r_over_l=r_over_l, a_seg=a_seg, a_blob=a_blob, **kw)


CILIA_WORKLOADS = {
    # 32 cilia x 20 segments on a 642-blob sphere (CPU-runnable)
    "cilia_1282": dict(nfil=32, nseg=20, nblob=642),
    # 512 cilia x 20 segments on a 4002-blob sphere
    "cilia_14242": dict(nfil=512, nseg=20, nblob=4002),
}


def cilia_workload(name, **kw):
    return cilia_config(**CILIA_WORKLOADS[name], **kw)


def carpet_seed(nfil, nseg, a_seg=1.0, seeding="rectangular"):
    """Filament bases of a carpet above the wall z = 0 (swimmer.cpp:68-165):
    a rectangular grid of spacing 1.4 L, or hexagonal rows 0.5 (1.0122 L +
    DL) apart, centred on the origin, bases 0.5 DL above the plane, with
    L = DL (NSEG - 1); filament ids run as the reference numbers them."""
    DL = 2.2 * a_seg
    L = DL * (nseg - 1)
    out = []
    if seeding == "rectangular":
        nx = int(math.sqrt(float(nfil)))
        ny = max(1, int(math.ceil(nfil / float(nx))))
        step_x = step_y = 1.4 * L
        im, jm = 0.5 * (nx - 1.0), 0.5 * (ny - 1.0)
        for i in range(nx):
            for j in range(ny):
                if j + i * ny < nfil:
                    out.append(((i - im) * step_x, (j - jm) * step_y, 0.5 * DL))
    elif seeding == "hexagonal":
        r3 = math.sqrt(3.0)
        nx = int(math.floor(0.25 * (3.0 - r3 + math.sqrt(4.0 - 2.0 * r3 + 8.0 * r3 * nfil)) + 0.5))
        if nx < 2:
            raise ValueError("hexagonal seeding needs rows of at least two filaments")
        ny = max(1, int(math.ceil(nfil / float(nx - 1))))
        step_y = 0.5 * (1.0122 * L + DL)
        step_x = 2.0 * step_y / r3
        im, jm = 0.5 * (nx - 1.0), 0.5 * (ny - 1.0)
        for j in range(ny):
            odd = j % 2
            for i in range(nx - odd):
                if len(out) < nfil:
                    out.append(((0.5 * odd + i - im) * step_x, (j - jm) * step_y, 0.5 * DL))
    else:
        raise ValueError(f"unknown seeding {seeding!r}")
    if len(out) != nfil:
        raise ValueError(f"the {seeding} grid holds {len(out)} of {nfil} filaments")
    return np.array(out, dtype=np.float64)


def carpet_config(nfil, nseg, a_seg=1.0, seeding="rectangular", **kw):
    """A carpet of prescribed cilia above a no-slip wall: the reference's
    INFINITE_PLANE_WALL build (config.hpp:276-289) -- one surface at rest, no
    blobs, RSEG = MU = 1 by default, phases 2 pi x / L + 2 pi 116/300 unless
    given (filament.cpp:222).  Phases are prescribed unless dynamic_phase and
    a generalised-force table are passed."""
    kw.setdefault("dynamic_phase", False)
    return CiliaConfig(nseg=nseg, nfil=np.array([nfil]), nblob=np.array([0]),
                       blob_refs=np.zeros((0, 3)), fil_refs=carpet_seed(nfil, nseg, a_seg, seeding),
                       body_x=np.zeros((1, 3)), a_seg=a_seg, fixed_body=True, wall=True, **kw)


CARPET_WORKLOADS = {
    # 64 cilia x 20 segments on a rectangular grid (CPU-runnable)
    "carpet_1280": dict(nfil=64, nseg=20),
    # 1024 cilia x 20 segments on a rectangular grid
    "carpet_20480": dict(nfil=1024, nseg=20),
}


def carpet_workload(name, **kw):
    return carpet_config(**CARPET_WORKLOADS[name], **kw)
