"""Learned warm start on the GPU: the hand-written forward/backward kernels
against the FP64 oracle (oracle/mlp.py), and what the trained network gains
over the extrapolation warm start on the 962-particle swimmer."""
import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry
from oracle import mlp

pytestmark = pytest.mark.gpu


def sim(**kw):
    cfg = geometry.sim_workload("step_962", seed=1, **kw)
    s = capi.Sim(cfg)
    s.make_body_reference_matrices()
    return cfg, s


def test_forward_and_gradient_match_oracle(gpu):
    cfg, s = sim()
    s.warmstart_init(hidden=128, seed=3)
    p = s.warmstart_params()
    rng = np.random.default_rng(1)
    X = rng.normal(size=(200, mlp.NIN)).astype(np.float32)
    T = rng.normal(size=(200, mlp.NOUT)).astype(np.float32)
    Y = s.warmstart_forward(X)
    Yo, _ = mlp.forward(p, X, 128)
    assert np.abs(Y - Yo).max() <= 1e-4 * max(1.0, np.abs(Yo).max())
    loss, g = s.warmstart_loss_grad(X, T)
    lo, go = mlp.loss_and_grad(p, X, T, 128)
    print(f"loss gpu {loss:.6e} oracle {lo:.6e}; grad rel err "
          f"{np.abs(g - go).max() / np.abs(go).max():.2e}", flush=True)
    assert abs(loss - lo) <= 1e-5 * lo
    assert np.abs(g - go).max() <= 1e-4 * np.abs(go).max()


def rollout(s, steps, nt0=0):
    its = []
    for nt in range(nt0, nt0 + steps):
        rc, bi, gi = s.step(nt)
        assert rc == 0, rc
        its.append((bi, gi))
    return np.array(its)


def test_trained_network_cuts_broyden_iterations(gpu):
    """Two identical swimmers run one driving period (300 steps) plainly; B
    records steps 5, 10, ..., 295 (59 x 320 segments = 18,880 rows; step 0
    has no history to learn from) and trains.  Over the next 150 steps
    A stays plain and B uses the network: B needs fewer Broyden iterations
    and both trajectories converge to the same tolerance."""
    cfg, a = sim()
    _, b = sim()
    b.warmstart_init(hidden=128, seed=7)
    b.warmstart_configure(apply=False, collect=20000, stride=5)
    rollout(a, 300)
    rollout(b, 300)
    rows = b.warmstart_dataset_size()
    assert rows == 59 * 320
    loss = b.warmstart_train(iters=3000, batch=1024, lr=1e-3, seed=1)
    print(f"dataset rows {rows}; training loss (standardised MSE) {loss:.4f}", flush=True)
    assert loss < 0.8
    b.warmstart_configure(apply=True, collect=0)
    ra = rollout(a, 150, 300)
    rb = rollout(b, 150, 300)
    applied, fallbacks = b.warmstart_stats()
    print(f"Broyden it/step plain {ra[:, 0].mean():.2f} learned {rb[:, 0].mean():.2f}; "
          f"GMRES it/step plain {ra[:, 1].mean():.1f} learned {rb[:, 1].mean():.1f}; "
          f"{applied} learned steps, {fallbacks} redone from the extrapolation", flush=True)
    assert applied == 150
    assert rb[:, 0].mean() < ra[:, 0].mean()
    sa, sb = a.state(), b.state()
    DL = cfg.swimmers[0].DL
    assert np.abs(sa["x"] - sb["x"]).max() <= 10 * DL * cfg.tol
