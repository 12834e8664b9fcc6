"""Parity at the benchmark workloads' own sizes (run on an MI355X with -m gpu).

Each test builds the system of one benchmark workload (mdt.workloads, the same seeds and
sizes as bench.py) and compares the HIP path with the FP64 oracle on it:
  C2  256,000 atoms, Q6.AVE + Q4.AVE + Lechner-Dellago local Q6, DIM v1+v2, WT, PACE + 1 steps
  C3  500,000-atom slab, S(k) over 2,056 k vectors, value and forces
  C4  64 walkers of 2,000-64,000 atoms, 2-D grids, every walker against the oracle
  C5  the C2 system with the 3-D CV (Q4.AVE, Q6.AVE, S(k_111)) on the exact hills sum

Bars: CV values and per-atom q_l 1e-5 relative; forces 1e-4 relative
(max |F - F_ref| / max |F_ref|); the selection's kept set and counts bit-exact; grids
1e-4 of their maximum; chain-rule weights 1e-4 of the natural scale h / sigma.
"""
import numpy as np
import pytest

from helpers import rel_force_err
from mdt import workloads as W

pytestmark = pytest.mark.gpu

CV_TOL = 1e-5
F_TOL = 1e-4


def _ctx(mdt, systems):
    ctx = mdt.Context(0, len(systems))
    for w, s in enumerate(systems):
        ctx.bind_system(w, s)
    return ctx


def _kept_reference(env, off, idx, n, N):
    """The oracle's kept neighbours of every atom in list order, -1 padded (n, N)."""
    pos = np.where(env.sel_pos >= 0, env.sel_pos, np.iinfo(np.int32).max).astype(np.int64)
    pos.sort(axis=1)
    ok = pos < np.iinfo(np.int32).max
    want = np.full((n, N), -1, np.int64)
    want[ok] = idx[(off[:n, None].astype(np.int64) + np.where(ok, pos, 0))[ok]]
    return want


def test_c2_full_size(mdt, orc):
    """The headline workload at 256,000 atoms for PACE + 1 steps (well-tempered deposits at
    steps 0 and 10), with two hills placed beside the CV so that the bias force is not zero."""
    s = W.c2_system(mdt)
    x, off, idx, own = s.arrays()
    n = s.n_local
    print(f"[c2] n={n} n_all={x.shape[0]} entries={off[-1]}", flush=True)
    env = orc.Env(x, off, idx, own, n, W.RC, 18)
    q6, q4 = env.q_ave(6), env.q_ave(4)
    _, ql6 = env.steinhardt(6)
    _, ql4 = env.steinhardt(4)
    _, lq = env.ld(6)
    lq6 = orc.kahan_mean(lq)
    G = env.dQ_dx(6, 1.0) + env.dQ_dx(4, 1.0)        # v1 + v2: both weights are dV/ds
    s_ref = q6 + q4
    ctx = _ctx(mdt, [s])
    fix = mdt.Fix(ctx, W.C2_ARGS)
    fix.init()
    hc, hh = np.array([[s_ref - 0.004], [s_ref + 0.011]]), np.array([0.005, 0.003])
    fix.add_hills(hc, hh, 0.01)
    g = orc.Grid([0.0], [2.0], [2000])
    for c, h in zip(hc, hh):
        g.deposit(c, h, 0.01)
    worst = cvw = 0.0
    heights = []
    for step in range(11):
        f = np.zeros((n, 3))
        fix.post_force(step, f)
        cvw = max(cvw, *(abs(fix.cv(k) - r) / r for k, r in (("Q6.AVE", q6), ("Q4.AVE", q4), ("LQ6.AVE", lq6))))
        assert abs(fix.cv("Q6.AVE") - q6) <= CV_TOL * q6, step
        assert abs(fix.cv("Q4.AVE") - q4) <= CV_TOL * q4, step
        assert abs(fix.cv("LQ6.AVE") - lq6) <= CV_TOL * lq6, step
        if step % 10 == 0:
            heights.append(g.wt_height([s_ref], 0.005, 10.0, 300.0, W.KB, 1))
            g.deposit([s_ref], heights[-1], 0.01)
        V, dV = g.eval([s_ref])
        assert abs(fix.scalar() - V) <= 1e-4 * abs(V), step
        w = fix.cv_weight("Q6.AVE")
        assert w == fix.cv_weight("Q4.AVE") and w != 0.0
        assert abs(w - dV[0]) <= 1e-4 * 0.005 / 0.01, (step, w, dV[0])
        e = rel_force_err(f, -w * G)
        worst = max(worst, e)
        assert e <= F_TOL, (step, e)
    print(f"[c2] worst CV rel err {cvw:.3e}, worst force rel err {worst:.3e}", flush=True)
    hills = fix.hills(1)             # the deposits of the run (added hills go to the grid only)
    assert list(hills[:, 0]) == [0.0, 10.0]
    assert np.abs(hills[:, 1] - s_ref).max() <= CV_TOL * s_ref
    assert np.abs(hills[:, 2] - heights).max() <= 1e-6 * max(heights)
    assert np.abs(fix.grid(2000) - g.g).max() <= 1e-4 * np.abs(g.g).max()
    for name, ref in (("Q6.stein_q", ql6), ("Q4.stein_q", ql4), ("LQ6.local_q", lq)):
        v = fix.extract(name, n)
        assert np.abs(v - ref).max() <= CV_TOL * np.abs(ref).max(), name
    assert np.array_equal(fix.extract("Q6.nsel", n).astype(np.int32), env.nsel)
    assert np.array_equal(fix.extract("Q6.count_rc", n).astype(np.int32), env.cnt)
    kept = fix.extract("Q6.kept", n, 18).astype(np.int64)
    assert np.array_equal(kept, _kept_reference(env, off, idx, n, 18))
    assert np.abs(f.sum(axis=0)).max() <= 1e-4 * np.abs(f).max() * np.sqrt(n)


def test_c3_full_size(mdt, orc):
    """500,000-atom slab, S(k) over the workload's 8 (111) + 2,048 seeded random k vectors.
    The grid spans the CV (the mean S over k is ~600 here) so the bias force is not zero."""
    s = W.c3_system(mdt)
    n = s.n_local
    x = s.arrays()[0][:n]
    print(f"[c3] n={n}", flush=True)
    ctx = _ctx(mdt, [s])
    args = W.C3_ARGS.replace("DIM 1 0 50 1000 v1", "DIM 1 0 2000 4000 v1")
    fix = mdt.Fix(ctx, args)
    k = fix.param("SK.KVEC").reshape(-1, 3)
    assert k.shape == (2056, 3)
    v_ref, A, S = orc.sk(x, k)
    print(f"[c3] S.AVE={v_ref:.6f}", flush=True)
    fix.init()
    fix.add_hills(np.array([[v_ref + 0.3]]), np.array([0.005]), 0.5)
    f = np.zeros((n, 3))
    fix.post_force(1, f)
    assert abs(fix.cv("SK.AVE") - v_ref) <= CV_TOL * abs(v_ref)
    print(f"[c3] CV rel err {abs(fix.cv('SK.AVE') - v_ref) / abs(v_ref):.3e}", flush=True)
    w = fix.cv_weight("SK.AVE")
    assert w != 0.0
    F_ref = -orc.sk_grad(x, k, w, A)
    e = rel_force_err(f, F_ref)
    print(f"[c3] force rel err {e:.3e}", flush=True)
    assert e <= F_TOL


def test_c4_full_size(mdt, orc):
    """64 walkers of uneven size in one context, each against the oracle: CVs, its own
    2-D grid (well-tempered deposit at step 0) and forces."""
    systems = W.c4_systems(mdt)
    nw = len(systems)
    ns = [sm.n_local for sm in systems]
    print(f"[c4] walkers={nw} atoms={sum(ns)}", flush=True)
    ctx = _ctx(mdt, systems)
    fix = mdt.Fix(ctx, W.C4_ARGS)
    fix.init()
    refs = []
    for w, sm in enumerate(systems):
        x, off, idx, own = sm.arrays()
        env = orc.Env(x, off, idx, own, sm.n_local, W.RC, 18)
        q4, q6 = env.q_ave(4), env.q_ave(6)
        hc = np.array([[q4 + 0.006, q6 - 0.004]])
        fix.add_hills(hc, np.array([0.005]), 0.01, walker=w)
        g = orc.Grid([0.0, 0.0], [1.0, 1.0], [200, 200])
        g.deposit(hc[0], 0.005, 0.01)
        refs.append((env, q4, q6, g))
    f = np.zeros((sum(ns), 3))
    fix.post_force(0, f)
    a = 0
    worst = cvw = 0.0
    for w, (env, q4, q6, g) in enumerate(refs):
        n = ns[w]
        cvw = max(cvw, abs(fix.cv("Q4.AVE", w) - q4) / q4, abs(fix.cv("Q6.AVE", w) - q6) / q6)
        assert abs(fix.cv("Q4.AVE", w) - q4) <= CV_TOL * q4, w
        assert abs(fix.cv("Q6.AVE", w) - q6) <= CV_TOL * q6, w
        s_ref = [q4, q6]
        g.deposit(s_ref, g.wt_height(s_ref, 0.005, 10.0, 300.0, W.KB, 1), 0.01)
        V, dV = g.eval(s_ref)
        w4, w6 = fix.cv_weight("Q4.AVE", w), fix.cv_weight("Q6.AVE", w)
        assert abs(w4 - dV[0]) <= 1e-4 * 0.005 / 0.01 and abs(w6 - dV[1]) <= 1e-4 * 0.005 / 0.01, w
        F_ref = -(env.dQ_dx(4, w4) + env.dQ_dx(6, w6))
        e = rel_force_err(f[a:a + n], F_ref)
        worst = max(worst, e)
        assert e <= F_TOL, (w, e)
        a += n
    print(f"[c4] worst CV rel err {cvw:.3e}, worst force rel err {worst:.3e}", flush=True)


def test_c5_full_size_exact_hills(mdt, orc):
    """The 3-D CV (Q4.AVE, Q6.AVE, S(k_111)) on the 256,000-atom system with the exact
    FP64 hills sum (Gaussian_Hill_type 2) over 2,000 well-tempered hills around it."""
    s = W.c2_system(mdt)
    x, off, idx, own = s.arrays()
    n = s.n_local
    print(f"[c5] n={n}", flush=True)
    ctx = _ctx(mdt, [s])
    fix = mdt.Fix(ctx, W.C5_BASE + "Gaussian_Hill_type 2")
    k = fix.param("S111.KVEC").reshape(-1, 3)
    env = orc.Env(x, off, idx, own, n, W.RC, 18)
    q4, q6 = env.q_ave(4), env.q_ave(6)
    v_sk, A, _ = orc.sk(x[:n], k)
    s_ref = np.array([q4, q6, v_sk])
    kt = 300.0 * W.KB
    hc, hh = W.wt_hills(s_ref, 2000, kt)
    fix.init()
    fix.add_hills(hc, hh, 0.01)
    f = np.zeros((n, 3))
    fix.post_force(1, f)                      # PACE 10: no deposit at step 1
    assert abs(fix.cv("Q4.AVE") - q4) <= CV_TOL * q4
    assert abs(fix.cv("Q6.AVE") - q6) <= CV_TOL * q6
    assert abs(fix.cv("S111.AVE") - v_sk) <= CV_TOL * v_sk
    rows = fix.hills(3)
    assert rows.shape[0] == 2000
    # the bias sum at the device's own CV point: S(k_111) is ~1e4 here, so its 1e-5
    # relative agreement is a large fraction of its hill width
    s_dev = np.array([fix.vector(d) for d in range(3)])
    assert np.abs(s_dev - s_ref).max() <= CV_TOL * np.abs(s_ref).max()
    print(f"[c5] s={s_ref} V at s_dev", flush=True)
    V, dV = orc.hills_sum(rows[:, 1:4], rows[:, 4], W.C5_SIGMAS, s_dev)
    assert abs(fix.scalar() - V[0]) <= 1e-9 * abs(V[0])
    wq4, wq6, ws = fix.cv_weight("Q4.AVE"), fix.cv_weight("Q6.AVE"), fix.cv_weight("S111.AVE")
    for wg, wr, sg in zip((wq4, wq6, ws), dV[0], W.C5_SIGMAS):
        assert abs(wg - wr) <= 1e-9 * np.abs(hh).max() / sg, (wg, wr)
    F_ref = -(env.dQ_dx(4, wq4) + env.dQ_dx(6, wq6) + orc.sk_grad(x[:n], k, ws, A))
    e = rel_force_err(f, F_ref)
    cvw = max(abs(fix.cv(k) - r) / r for k, r in (("Q4.AVE", q4), ("Q6.AVE", q6), ("S111.AVE", v_sk)))
    print(f"[c5] CV rel err {cvw:.3e}, force rel err {e:.3e}, V rel err {abs(fix.scalar() - V[0]) / abs(V[0]):.3e}",
          flush=True)
    assert e <= F_TOL
