"""GPU parity of the gfx950 pair kernels against the CPU FP64 oracle.

Bars: FP64 kernels within 1e-12 (max-norm relative per output block) of the
oracle; FP32-Kahan within 1e-6 of the FP64 result.  Every test goes through
the C ABI (libgmi.so) only.
"""
import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry
from oracle import binding as orc

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a); b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a).max())


def run_blocks(ctx, xs, xb, fs, fb):
    """Each reference block separately plus the fused grand apply."""
    ctx.upload(xs, fs, xb, fb)
    out = {}
    ctx.call("gmi_evaluate_segment_segment_mobility")
    out["ss"] = ctx.velocities()[0].copy()
    ctx.call("gmi_evaluate_segment_blob_mobility")      # += onto V_s
    out["ss+sb"] = ctx.velocities()[0].copy()
    ctx.call("gmi_evaluate_blob_segment_mobility")
    out["bs"] = ctx.velocities()[1].copy()
    ctx.call("gmi_evaluate_blob_blob_mobility")
    out["bb"] = ctx.velocities()[1].copy()
    ctx.call("gmi_evaluate_grand_mobility")
    out["grand"] = ctx.velocities()
    return out


def oracle_blocks(xs, xb, fs, fb, a_s, a_b):
    ss = orc.Mss(xs, fs, a_s)
    return {"ss": ss, "ss+sb": orc.Msb(xs, xb, fb, a_s, a_b, V=ss),
            "bs": orc.Mbs(xb, xs, fs, a_s, a_b), "bb": orc.Mbb(xb, fb, a_b)}


def check_fp64(out, ref):
    for k in ("ss", "ss+sb", "bs", "bb"):
        e = relerr(out[k], ref[k])
        assert e <= 1e-12, (k, e)
    vs, vb = out["grand"]
    assert relerr(vs, ref["ss+sb"]) <= 1e-12
    assert relerr(vb, ref["bs"] + ref["bb"]) <= 1e-12


def test_apply_962_fp64_parity(gpu):
    g = geometry.workload_geometry("apply_962")
    xs, xb = g.positions()
    fs, fb = geometry.random_forces(g.ns, g.nblob, seed=1)
    print(f"apply_962 fp64: Ns={g.ns} Nb={g.nblob}", flush=True)
    ctx = capi.Context([g.ns], [g.nblob], g.nseg)
    check_fp64(run_blocks(ctx, xs, xb, fs, fb), oracle_blocks(xs, xb, fs, fb, 1.0, 1.0))


@pytest.mark.parametrize("ns,nb,nseg,a_b,seed", [
    (40, 37, 4, 1.4, 2),      # unequal radii, ragged tiles
    (3, 1, 3, 0.6, 3),        # tiny
    (600, 300, 20, 1.4, 4),   # several tiles, splits
    (20, 0, 20, 1.0, 5),      # no blobs
    (0, 50, 1, 1.2, 6),       # no segments
])
def test_random_cluster_fp64_parity(gpu, ns, nb, nseg, a_b, seed):
    rng = np.random.default_rng(seed)
    L = max(2.0, (ns + nb) ** (1 / 3) * 1.6)    # dense: many overlaps
    xs = rng.uniform(-L, L, (ns, 3)); xb = rng.uniform(-L, L, (nb, 3))
    if ns and nb:
        xb[0] = xs[0]                              # coincident centres
    fs = rng.normal(size=(ns, 6)); fb = rng.normal(size=(nb, 3))
    print(f"cluster ns={ns} nb={nb} a_b={a_b}", flush=True)
    ctx = capi.Context([ns], [nb], nseg, a_blob=a_b)
    out = run_blocks(ctx, xs, xb, fs, fb)
    ref = oracle_blocks(xs, xb, fs, fb, 1.0, a_b)
    if ns and nb:
        check_fp64(out, ref)
    elif ns:
        assert relerr(out["ss"], ref["ss"]) <= 1e-12
        assert np.array_equal(out["ss+sb"], out["ss"])
    else:
        assert relerr(out["bb"], ref["bb"]) <= 1e-12
        assert np.all(out["bs"] == 0)


def test_apply_962_fp32_kahan_error(gpu):
    g = geometry.workload_geometry("apply_962")
    xs, xb = g.positions()
    fs, fb = geometry.random_forces(g.ns, g.nblob, seed=1)
    ctx = capi.Context([g.ns], [g.nblob], g.nseg, precision=capi.GMI_FP32_KAHAN)
    ctx.upload(xs, fs, xb, fb)
    ctx.call("gmi_evaluate_grand_mobility")
    vs, vb = ctx.velocities()
    ref = oracle_blocks(xs, xb, fs, fb, 1.0, 1.0)
    es, eb = relerr(vs, ref["ss+sb"]), relerr(vb, ref["bs"] + ref["bb"])
    print(f"fp32 kahan rel err: segs {es:.2e} blobs {eb:.2e}", flush=True)
    assert es <= 1e-6 and eb <= 1e-6


def test_independent_swimmers_are_block_diagonal(gpu):
    rng = np.random.default_rng(7)
    sizes = [(40, 30), (60, 0), (20, 90)]
    xs_l, xb_l, fs_l, fb_l = [], [], [], []
    for ns, nb in sizes:
        c = rng.normal(size=3) * 3        # overlapping swimmers in space
        xs_l.append(c + rng.uniform(-4, 4, (ns, 3))); xb_l.append(c + rng.uniform(-4, 4, (nb, 3)))
        fs_l.append(rng.normal(size=(ns, 6))); fb_l.append(rng.normal(size=(nb, 3)))
    ctx = capi.Context([s[0] for s in sizes], [s[1] for s in sizes], 20, a_blob=1.3,
                       interact=False)
    ctx.upload(np.vstack(xs_l), np.vstack(fs_l), np.vstack(xb_l), np.vstack(fb_l))
    ctx.call("gmi_evaluate_grand_mobility")
    vs, vb = ctx.velocities()
    s0 = b0 = 0
    for (ns, nb), xs, xb, fs, fb in zip(sizes, xs_l, xb_l, fs_l, fb_l):
        ref = oracle_blocks(xs, xb, fs, fb, 1.0, 1.3)
        if ns:
            assert relerr(vs[s0:s0 + ns], ref["ss+sb"]) <= 1e-12
        if nb:
            assert relerr(vb[b0:b0 + nb], ref["bs"] + ref["bb"]) <= 1e-12
        s0 += ns; b0 += nb
    # interacting: the same particles all couple (reference NSWIM semantics)
    ctx2 = capi.Context([s[0] for s in sizes], [s[1] for s in sizes], 20, a_blob=1.3,
                        interact=True)
    ctx2.upload(np.vstack(xs_l), np.vstack(fs_l), np.vstack(xb_l), np.vstack(fb_l))
    ctx2.call("gmi_evaluate_grand_mobility")
    vs2, vb2 = ctx2.velocities()
    ref = oracle_blocks(np.vstack(xs_l), np.vstack(xb_l), np.vstack(fs_l), np.vstack(fb_l),
                        1.0, 1.3)
    assert relerr(vs2, ref["ss+sb"]) <= 1e-12
    assert relerr(vb2, ref["bs"] + ref["bb"]) <= 1e-12


def test_barrier_parity(gpu):
    rng = np.random.default_rng(8)
    nseg, nfil, nb = 5, 8, 70
    xs = rng.uniform(-4, 4, (nseg * nfil, 3)); xb = rng.uniform(-4, 4, (nb, 3))
    fs = rng.normal(size=(nseg * nfil, 6))
    K = 3.5
    ctx = capi.Context([nseg * nfil], [nb], nseg, a_blob=1.2)
    ctx.upload(xs, fs, xb, np.zeros((nb, 3)))
    ctx.call("gmi_apply_interparticle_forces", K)
    rep = np.zeros(3 * nb)
    ctx.call("gmi_copy_interparticle_blob_forces_to_host", capi.ptr(rep))
    ctx.call("gmi_wait_for_device")
    # the barrier acts on the device segment forces: read them back through
    # an ss apply is indirect, so compare M_ss (F + barrier) instead
    ctx.call("gmi_evaluate_segment_segment_mobility")
    vs, _ = ctx.velocities()
    f_ref, rep_ref = orc.barrier(xs, xb, fs, nseg, 1.0, 1.2, K)
    assert relerr(rep.reshape(-1, 3), rep_ref) <= 1e-13
    assert relerr(vs, orc.Mss(xs, f_ref, 1.0)) <= 1e-12


@pytest.mark.parametrize("interact", [True, False])
def test_barrier_two_swimmers_parity(gpu, interact):
    """Two swimmers filling the same region: coupled, every particle pair
    counts; independent, only pairs inside one swimmer (the cell grid keys
    buckets by swimmer).  Against the oracle's barrier over the union, or
    over each swimmer alone."""
    rng = np.random.default_rng(11)
    nseg, nfil, nb = 6, (7, 5), (60, 45)
    K = 2.5
    xs = [rng.uniform(-4, 4, (nseg * nf, 3)) for nf in nfil]
    xb = [rng.uniform(-4, 4, (n, 3)) for n in nb]
    fs = [rng.normal(size=(nseg * nf, 6)) for nf in nfil]
    ctx = capi.Context([nseg * nf for nf in nfil], list(nb), nseg, a_blob=1.2, interact=interact)
    ctx.upload(np.vstack(xs), np.vstack(fs), np.vstack(xb), np.zeros((sum(nb), 3)))
    ctx.call("gmi_apply_interparticle_forces", K)
    rep = np.zeros(3 * sum(nb))
    ctx.call("gmi_copy_interparticle_blob_forces_to_host", capi.ptr(rep))
    ctx.call("gmi_wait_for_device")
    ctx.call("gmi_evaluate_segment_segment_mobility")
    vs, _ = ctx.velocities()
    if interact:
        f_ref, rep_ref = orc.barrier(np.vstack(xs), np.vstack(xb), np.vstack(fs), nseg, 1.0, 1.2, K)
        v_ref = orc.Mss(np.vstack(xs), f_ref, 1.0)
    else:
        parts = [orc.barrier(xs[k], xb[k], fs[k], nseg, 1.0, 1.2, K) for k in range(2)]
        rep_ref = np.vstack([p[1] for p in parts])
        v_ref = np.vstack([orc.Mss(xs[k], parts[k][0], 1.0) for k in range(2)])
    assert np.abs(rep_ref).max() > 0
    assert relerr(rep.reshape(-1, 3), rep_ref) <= 1e-13
    assert relerr(vs, v_ref) <= 1e-12


@pytest.mark.parametrize("layout", ["crowded", "far"])
def test_barrier_cell_grid_edge_cases(gpu, layout):
    """The barrier's cell grid on inputs that reach its rare paths: many
    particles in a few cells (buckets longer than the in-register sort), and
    a cluster far from the origin on the negative side (large negative cell
    coordinates in the hash and the packed keys)."""
    rng = np.random.default_rng(5)
    nseg, nfil, nb = 4, 40, 120
    if layout == "crowded":
        xs = rng.uniform(-1.2, 1.2, (nseg * nfil, 3)); xb = rng.uniform(-1.2, 1.2, (nb, 3))
    else:
        off = np.array([-9.5e3, 4.1e3, -7.7e3])
        xs = off + rng.uniform(-6, 6, (nseg * nfil, 3)); xb = off + rng.uniform(-6, 6, (nb, 3))
    fs = rng.normal(size=(nseg * nfil, 6))
    K = 1.7
    ctx = capi.Context([nseg * nfil], [nb], nseg, a_blob=0.8)
    ctx.upload(xs, fs, xb, np.zeros((nb, 3)))
    ctx.call("gmi_apply_interparticle_forces", K)
    rep = np.zeros(3 * nb)
    ctx.call("gmi_copy_interparticle_blob_forces_to_host", capi.ptr(rep))
    ctx.call("gmi_wait_for_device")
    ctx.call("gmi_evaluate_segment_segment_mobility")
    vs, _ = ctx.velocities()
    f_ref, rep_ref = orc.barrier(xs, xb, fs, nseg, 1.0, 0.8, K)
    assert np.abs(rep_ref).max() > 0
    assert relerr(rep.reshape(-1, 3), rep_ref) <= 1e-13
    assert relerr(vs, orc.Mss(xs, f_ref, 1.0)) <= 1e-12


_REF_46722 = {}


def _oracle_46722(g, xs, xb, fs, fb):
    """The oracle's whole 46,722-particle grand apply (all rows; about a
    second on 16 host threads), computed once per session."""
    if not _REF_46722:
        _REF_46722["v"] = orc.grand(xs, xb, fs, fb, 1.0, 1.0)
    return _REF_46722["v"]


@pytest.mark.parametrize("prec", [capi.GMI_FP64, capi.GMI_FP32_KAHAN])
def test_apply_46722_all_rows(gpu, prec):
    """Full 46,722-particle apply on the GPU against the oracle on every row
    (FP64: 1e-12; FP32-Kahan: 1e-6), relative to the largest entry of each
    species' velocities."""
    g = geometry.workload_geometry("apply_46722")
    xs, xb = g.positions()
    fs, fb = geometry.random_forces(g.ns, g.nblob, seed=3)
    print(f"apply_46722 prec={prec}", flush=True)
    ctx = capi.Context([g.ns], [g.nblob], g.nseg, precision=prec)
    ctx.upload(xs, fs, xb, fb)
    ctx.call("gmi_evaluate_grand_mobility")
    vs, vb = ctx.velocities()
    ctx.close()
    rs, rb = _oracle_46722(g, xs, xb, fs, fb)
    tol = 1e-12 if prec == capi.GMI_FP64 else 1e-6
    es, eb = relerr(vs, rs), relerr(vb, rb)
    print(f"  all rows: segments {es:.2e}, blobs {eb:.2e}", flush=True)
    assert es <= tol and eb <= tol
