"""Learned bias on the GPU: MFMA kernels against the FP64 oracle, fit, held-out gate, fallback.

Tolerances (written per test):
  forward vs FP64 network            5e-4 x max|V|   (BF16 hi+lo operands, FP32 tanh/accumulation)
  gradient vs FP64 autograd oracle   cosine >= 0.999, relative L2 error <= 5e-2
  fitted bias vs exact hills sum     max |dV| <= 0.1 kT on the held-out set (the gate) and
                                     <= 0.15 kT on fresh points
"""
import time

import numpy as np
import pytest

from helpers import fcc_system

pytestmark = pytest.mark.gpu

KT = 300.0 * 8.617343e-5
H = 64


def _bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _split(p):
    p = np.asarray(p, np.float64)
    W1 = p[0:192].reshape(H, 3)
    b1 = p[192:256]
    W2 = p[256:256 + 4096].reshape(H, H)
    b2 = p[4352:4416]
    W3 = p[4416:4480]
    return W1, b1, W2, b2, W3, p[4480]


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(4481, np.float32)
    p[0:192] = rng.normal(size=192) * 0.8
    p[192:256] = rng.normal(size=64) * 0.2
    p[256:4352] = rng.normal(size=4096) / 8
    p[4352:4416] = rng.normal(size=64) * 0.1
    p[4416:4480] = rng.normal(size=64) / 8
    p[4480] = 0.05
    return p


def _ctx(mdt):
    return mdt.Context(0, 1)


def test_mlp_forward_matches_oracle(mdt, orc):
    ctx = _ctx(mdt)
    p = _params(1)
    norm = np.array([0.5, 0.4, 100.0, 8.0, 6.0, 0.5, 0.3, 0.01])
    rng = np.random.default_rng(2)
    s = np.column_stack([rng.uniform(0.4, 0.6, 1000), rng.uniform(0.3, 0.5, 1000), rng.uniform(98, 102, 1000)])
    V = mdt.mlp_forward(ctx, p, norm, s)
    x = (s.astype(np.float32).astype(np.float64) - norm[:3]) * norm[3:6]
    W1, b1, W2, b2, W3, b3 = _split(p)
    Vref, _ = orc.mlp_forward([W1, b1, W2, b2, W3, b3], x)
    Vref = Vref * norm[6] + norm[7]
    scale = np.abs(Vref).max()
    assert np.abs(V - Vref).max() <= 5e-4 * scale


def test_mlp_gradient_matches_oracle(mdt, orc):
    ctx = _ctx(mdt)
    p = _params(3)
    norm = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.7, 0.0])
    rng = np.random.default_rng(4)
    s = rng.uniform(-1, 1, size=(777, 3))        # not a multiple of the 64-point block
    v = rng.uniform(-0.5, 0.5, size=777)
    g = mdt.mlp_loss_grad(ctx, p, norm, s, v).astype(np.float64)
    x = s.astype(np.float32).astype(np.float64)
    W1, b1, W2, b2, W3, b3 = _split(p)
    _, gref = orc.mlp_loss_grad([W1, b1, W2, b2, W3, b3], x, v / norm[6])
    parts = [(0, 192), (192, 256), (256, 4352), (4352, 4416), (4416, 4480)]
    for (a, b), gr in zip(parts, gref[:5]):
        gg, rr = g[a:b], np.asarray(gr).ravel()
        cos = gg @ rr / (np.linalg.norm(gg) * np.linalg.norm(rr))
        assert cos >= 0.999, (a, cos)
        assert np.linalg.norm(gg - rr) <= 5e-2 * np.linalg.norm(rr), (a, np.linalg.norm(gg - rr) / np.linalg.norm(rr))
    assert abs(g[4480] - gref[5]) <= 5e-2 * abs(gref[5]) + 1e-6


def _hill_path(n, seed, h0=0.003):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    c = np.column_stack([0.30 + 0.03 * np.sin(t / 40.0) + 0.002 * rng.normal(size=n),
                         0.50 + 0.00008 * t + 0.002 * rng.normal(size=n),
                         100.0 + 1.5 * np.cos(t / 55.0) + 0.1 * rng.normal(size=n)])
    h = h0 * np.exp(-t / 500.0)
    return c, h


def _hill_cloud(n, seed, h0=0.0008):
    """Hills of a run trapped in one basin: a cloud around one CV point."""
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    c = np.array([0.3, 0.5, 100.0]) + rng.normal(size=(n, 3)) * np.array([0.007, 0.007, 0.35])
    return c, h0 * np.exp(-t / 500.0)


FIX3 = ("GAUSSIAN 0.01 0.003 10 PACE 0 HILLS_FILE NONE TEMP 300 Gaussian_Hill_type 3 "
        "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
        "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
        "CAL NAME LQ6 STEINH LD 6 all cutoff_r 3.9 cutoff_Natoms 18 "
        "CV_dim 3 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE SYMBOL v3 LQ6.AVE "
        "DIM 1 0 1 100 v1 DIM 2 0 1 100 v2 DIM 3 0 200 100 v3 SIGMAS 0.01 0.01 0.5 WT 0 ")


def test_mlp_fit_gate_and_recent_hills(mdt, orc):
    s = fcc_system(mdt, 5, sigma=0.1, seed=3)
    ctx = _ctx(mdt)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, FIX3 + "MLP_POINTS 200000 MLP_HOLDOUT 20000 MLP_EPOCHS_FIRST 200 MLP_BATCH 4096")
    fix.init()
    c, h = _hill_cloud(150, 7)
    fix.add_hills(c, h, 0.01)
    t0 = time.perf_counter()
    fix.mlp_refit()
    st = fix.mlp_status()
    print(f"fit: {st} in {time.perf_counter() - t0:.3f} s; tol {0.1 * KT:.3e}", flush=True)
    assert st["fits"] == 1 and st["n_fit"] == 150
    assert st["ok"] == (1 if st["err"] <= 0.1 * KT else 0)
    assert st["ok"] == 1, st
    sig = np.array([0.01, 0.01, 0.5])
    rng = np.random.default_rng(11)
    pts = c[rng.integers(0, 150, 2000)] + rng.normal(size=(2000, 3)) * sig
    V, dV = fix.bias_eval(pts)
    Vx, dVx = orc.hills_sum(c, h, sig, pts)
    assert np.abs(V - Vx).max() <= 0.15 * KT
    # hills deposited after the fit are summed exactly on top of the network
    c2, h2 = _hill_cloud(160, 8)
    fix.add_hills(c2[150:], h2[150:] * 3, 0.01)
    V2, _ = fix.bias_eval(pts)
    call, hall = np.vstack([c, c2[150:]]), np.concatenate([h, h2[150:] * 3])
    Vx2, _ = orc.hills_sum(call, hall, sig, pts)
    assert np.abs(V2 - Vx2).max() <= 0.15 * KT
    assert np.abs((V2 - V) - (Vx2 - Vx)).max() <= 1e-9 * max(1.0, np.abs(Vx2 - Vx).max())
    # the learned gradient is close to the exact one where the bias matters
    big = np.abs(dVx).max(axis=0)
    assert np.all(np.abs(dV - dVx).max(axis=0) <= 0.1 * big)
    # a refit keeps its training points and adds the ten new hills to their labels (and
    # redraws a block around them); its held-out points are labelled afresh by every hill,
    # so stale labels would show as a held-out error of the new hills' size
    fix.mlp_refit()
    st2 = fix.mlp_status()
    print(f"refit: {st2}", flush=True)
    assert st2["fits"] == 2 and st2["n_fit"] == 160
    assert st2["err"] <= 0.1 * KT, st2
    V3, _ = fix.bias_eval(pts)
    assert np.abs(V3 - Vx2).max() <= 0.15 * KT


def test_mlp_fallback_is_the_exact_sum(mdt, orc):
    s = fcc_system(mdt, 5, sigma=0.1, seed=3)
    ctx = _ctx(mdt)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, FIX3 + "MLP_POINTS 20000 MLP_HOLDOUT 4096 MLP_EPOCHS_FIRST 1 MLP_TOL 1e-9")
    fix.init()
    c, h = _hill_path(120, 5)
    fix.add_hills(c, h, 0.01)
    fix.mlp_refit()
    st = fix.mlp_status()
    assert st["ok"] == 0 and st["fallbacks"] == 1 and st["err"] > 1e-9 * KT
    sig = np.array([0.01, 0.01, 0.5])
    pts = c[::7] + 0.003
    V, dV = fix.bias_eval(pts)
    Vx, dVx = orc.hills_sum(c, h, sig, pts)
    assert np.abs(V - Vx).max() <= 1e-12 * np.abs(Vx).max()
    assert np.abs(dV - dVx).max() <= 1e-10 * np.abs(dVx).max()


def test_mlp_refits_inside_a_run(mdt, orc):
    s = fcc_system(mdt, 5, sigma=0.1, seed=9)
    n = s.n_local
    ctx = _ctx(mdt)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, FIX3.replace("PACE 0", "PACE 2") +
                  "MLP_REFIT 10 MLP_POINTS 30000 MLP_HOLDOUT 4096 MLP_EPOCHS_FIRST 20 MLP_EPOCHS 4 MLP_BATCH 4096")
    x, off, idx, own = s.arrays()
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    for step in range(30):
        f = np.zeros((n, 3))
        fix.post_force(step, f)
    st = fix.mlp_status()
    assert st["fits"] == 3 and st["n_fit"] == 15
    hills = fix.hills(3)
    assert hills.shape[0] == 15
    # step 30: V from the network (+ recent hills) against the exact sum at the current CV
    f = np.zeros((n, 3))
    fix.post_force(30, f)
    sv = np.array([fix.vector(k) for k in range(3)])
    hills = fix.hills(3)
    Vx, dVx = orc.hills_sum(hills[:, 1:4], hills[:, 4], [0.01, 0.01, 0.5], sv[None, :])
    if fix.mlp_status()["ok"]:
        assert abs(fix.scalar() - Vx[0]) <= 0.15 * KT
    else:
        assert abs(fix.scalar() - Vx[0]) <= 1e-12 * abs(Vx[0])
    w4, w6 = fix.cv_weight("Q4.AVE"), fix.cv_weight("Q6.AVE")
    wl = fix.cv_weight("LQ6.AVE")
    F_ref = -(env.dQ_dx(4, w4) + env.dQ_dx(6, w6) + env.dLQ_dx(6, wl))
    assert np.abs(f - F_ref).max() <= 1e-4 * np.abs(F_ref).max()


def test_mlp_gain_over_exact_sum(mdt, orc):
    """What the network gains: evaluating the bias at many points (FES maps, reweighting,
    batched walkers) costs a fixed 8.7 kFLOP per point instead of a sweep over every hill."""
    s = fcc_system(mdt, 5, sigma=0.1, seed=3)
    ctx = _ctx(mdt)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, FIX3.replace("Gaussian_Hill_type 3", "Gaussian_Hill_type 2"))
    fix.init()
    c, h = _hill_path(20000, 1)
    fix.add_hills(c, h, 0.01)
    rng = np.random.default_rng(0)
    pts = c[rng.integers(0, 20000, 20000)] + rng.normal(size=(20000, 3)) * np.array([0.01, 0.01, 0.5])
    fix.bias_eval(pts[:64])
    t0 = time.perf_counter()
    fix.bias_eval(pts)
    t_exact = time.perf_counter() - t0
    p = _params(5)
    norm = np.array([0.3, 0.5, 100.0, 20.0, 1.0, 0.5, 0.1, 0.0])
    mdt.mlp_forward(ctx, p, norm, pts[:64])
    t0 = time.perf_counter()
    mdt.mlp_forward(ctx, p, norm, pts)
    t_mlp = time.perf_counter() - t0
    print(f"bias at 20,000 points with 20,000 hills: exact {t_exact * 1e3:.2f} ms, network {t_mlp * 1e3:.2f} ms",
          flush=True)
    assert t_mlp < t_exact


def test_mlp_refits_in_batched_runs(mdt):
    """A run of 23 steps with refits every 10 gives the same bits with one step per
    graph launch, with 8, and with plain kernel launches: the launches stop at each
    refit step, so the fits see the same hills and the steps after them the same
    network."""
    s = fcc_system(mdt, 5, sigma=0.1, seed=9)
    n = s.n_local
    args = (FIX3.replace("PACE 0", "PACE 2") +
            "MLP_REFIT 10 MLP_POINTS 30000 MLP_HOLDOUT 4096 MLP_EPOCHS_FIRST 20 MLP_EPOCHS 4 MLP_BATCH 4096")
    out = []
    for per_launch in (1, 8, 0):
        ctx = _ctx(mdt)
        ctx.bind_system(0, s)
        fix = mdt.Fix(ctx, args)
        fix.init()
        fix.zero_forces()
        fix.run(0, 23, per_launch)
        fix.sync()
        st = fix.mlp_status()
        out.append((fix.forces(n), fix.scalar(), fix.hills(3), st["fits"], st["n_fit"], fix.mlp_params()))
        print(f"[mlp batched] per_launch={per_launch} fits={st['fits']} n_fit={st['n_fit']} V={out[-1][1]:.6g}",
              flush=True)
    a = out[0]
    for b in out[1:]:
        assert a[3] == b[3] == 2 and a[4] == b[4] == 10
        assert np.array_equal(a[2], b[2]) and a[2].shape[0] == 12
        assert all(np.array_equal(u, v) for u, v in zip(a[5], b[5]))
        assert np.array_equal(a[0], b[0]) and a[1] == b[1]
