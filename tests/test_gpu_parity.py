"""GPU parity of the HIP path against the FP64 oracle (run on an MI355X with -m gpu).

Bars (written in each test):
  selection (kept set of the step, and slot order)  bit-exact
  per-atom q_l and CV values          1e-5 relative
  forces                              1e-4 relative (max |F - F_ref| / max |F_ref|)
"""
import os

import numpy as np
import pytest

from helpers import Mover, fcc_system, liquid_system, q46_fix, rel_force_err

pytestmark = pytest.mark.gpu

CV_TOL = 1e-5
F_TOL = 1e-4


def _ctx(mdt, system):
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, system)
    return ctx


def _fix_q(mdt, ctx, N=18, rc=3.9, extra=""):
    return mdt.Fix(ctx, q46_fix(extra=extra, N=N, rc=rc))


@pytest.mark.parametrize("case", ["fcc_jitter", "fcc_perfect", "liquid_capped", "fcc_at_cutoff", "far_box"])
def test_selection_bit_exact(mdt, orc, case):
    if case == "fcc_jitter":
        s, N, rc = fcc_system(mdt, 10, sigma=0.1, seed=11), 18, 3.9
    elif case == "fcc_perfect":
        s, N, rc = fcc_system(mdt, 6, sigma=0.0), 8, 3.9          # ties at the cap
    elif case == "fcc_at_cutoff":
        # rc equal to the nearest-neighbour distance: every first-shell r^2 sits on rc^2
        # within rounding, so the in/out decisions are those of the exact FP64 r^2
        s, N, rc = fcc_system(mdt, 6, sigma=0.0), 18, 5.26 / np.sqrt(2.0)
    elif case == "far_box":
        # coordinates near 5,000 A: the FP32 rounding margin of the list scan is wide
        L = 8 * 5.26
        s = mdt.System().set_box([5000.0] * 3, [5000.0 + L] * 3).add_fcc(8, 8, 8, 5.26, 0.1, 0.1, 0.0, seed=13)
        s.build(5.9)
        N, rc = 18, 3.9
    else:
        s, N, rc = liquid_system(mdt, 3000, density=0.05, seed=5), 12, 3.9   # cap binds often
    print(f"[selection] case={case} n={s.n_local} N={N}", flush=True)
    x, off, idx, own = s.arrays()
    n = s.n_local
    ctx = _ctx(mdt, s)
    fix = _fix_q(mdt, ctx, N=N, rc=rc)
    fix.post_force(0)
    env = orc.Env(x, off, idx, own, n, rc, N)
    # the step's own selection: the set its k_env kept (decoded from the step's mask),
    # in list order, against the oracle's literal insertion
    kept = fix.extract("Q6.kept", n, N).astype(np.int64)
    want = np.full((n, N), -1, np.int64)
    for i in range(n):
        pos = np.sort(env.sel_pos[i][env.sel_pos[i] >= 0])
        want[i, : pos.size] = idx[off[i] + pos]
    assert np.array_equal(kept, want), f"kept set differs at {np.nonzero((kept != want).any(axis=1))[0][:5]}"
    assert np.array_equal(fix.extract("Q6.nsel", n).astype(np.int32), env.nsel)
    assert np.array_equal(fix.extract("Q6.count_rc", n).astype(np.int32), env.cnt)
    # slot order: the export instance of the same kernel runs the insertion for every atom
    sel, cnt, ns = fix.selection("Q6", n, N)
    assert np.array_equal(cnt, env.cnt)
    assert np.array_equal(ns, env.nsel)
    assert np.array_equal(sel, env.sel_idx), "slot order differs"
    if case == "liquid_capped":
        assert (env.cnt > N).sum() > 100
    if case == "fcc_at_cutoff":
        assert env.cnt.min() < env.cnt.max() <= 12     # the boundary splits the first shell


def test_q_values_and_forces_c1(mdt, orc):
    """4,000-atom jittered FCC (the CPU reference configuration), Q6+Q4 on v1+v2."""
    s = fcc_system(mdt, 10, sigma=0.1, seed=1)
    x, off, idx, own = s.arrays()
    n = s.n_local
    print(f"[q/forces] n={n}", flush=True)
    ctx = _ctx(mdt, s)
    fix = _fix_q(mdt, ctx)
    # put hills around the current CV so dV/ds is nonzero
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    s0 = env.q_ave(6) + env.q_ave(4)
    fix.init()
    fix.add_hills(np.array([[s0 - 0.004], [s0 + 0.011]]), np.array([0.005, 0.003]), 0.01)
    f = np.zeros((n, 3))
    fix.post_force(1, f)     # step 1: no deposit with PACE 10
    for l, name in ((6, "Q6"), (4, "Q4")):
        q_ref = env.steinhardt(l)[1]
        q = fix.extract(f"{name}.stein_q", n)
        assert np.abs(q - q_ref).max() <= CV_TOL * np.abs(q_ref).max()
        assert abs(fix.cv(f"{name}.AVE") - env.q_ave(l)) <= CV_TOL * env.q_ave(l)
    w6, w4 = fix.cv_weight("Q6.AVE"), fix.cv_weight("Q4.AVE")
    assert w6 != 0.0 and w6 == w4
    F_ref = -(env.dQ_dx(6, w6) + env.dQ_dx(4, w4))
    err = rel_force_err(f, F_ref)
    print(f"force rel err {err:.3e}")
    assert err <= F_TOL
    # translation invariance of the exact gradient
    assert np.abs(f.sum(axis=0)).max() <= 1e-4 * np.abs(f).max() * np.sqrt(n)
    # virial of the bias forces, the reference tally sum F (x) unwrap(x)
    img = np.random.default_rng(5).integers(-2, 3, size=(n, 3)).astype(np.int32)
    L = s.box()[1] - s.box()[0]
    u = x[:n] + img * L
    ref = np.array([f[:, 0] @ u[:, 0], f[:, 1] @ u[:, 1], f[:, 2] @ u[:, 2],
                    f[:, 0] @ u[:, 1], f[:, 0] @ u[:, 2], f[:, 1] @ u[:, 2]])
    vir = fix.virial(0, img)
    assert np.abs(vir - ref).max() <= 1e-9 * np.abs(ref).max()


def test_forces_when_cap_binds(mdt, orc):
    """Asymmetric selections (N cap binds): reverse pairs must still be exact."""
    s = liquid_system(mdt, 3000, density=0.05, seed=9)
    x, off, idx, own = s.arrays()
    n = s.n_local
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, q46_fix(N=12, lo=0.0, hi=1.5, nb=1500))
    env = orc.Env(x, off, idx, own, n, 3.9, 12)
    s0 = env.q_ave(6) + env.q_ave(4)
    fix.init()
    fix.add_hills(np.array([[s0 + 0.006]]), np.array([0.005]), 0.01)
    f = np.zeros((n, 3))
    fix.post_force(1, f)
    w = fix.cv_weight("Q6.AVE")
    F_ref = -(env.dQ_dx(6, w) + env.dQ_dx(4, w))
    assert rel_force_err(f, F_ref) <= F_TOL


def test_ld_local_q(mdt, orc):
    """Lechner-Dellago local Q6 value and forces (biased alone)."""
    s = fcc_system(mdt, 8, sigma=0.3, seed=4)
    x, off, idx, own = s.arrays()
    n = s.n_local
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE NONE "
                       "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                       "CAL NAME LQ6 STEINH LD 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                       "CV_dim 1 SYMBOL v1 LQ6.AVE DIM 1 0 1 1000 v1 WT 0")
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    qb, lq = env.ld(6)
    fix.init()
    fix.add_hills(np.array([[lq.mean() - 0.005]]), np.array([0.005]), 0.01)
    f = np.zeros((n, 3))
    fix.post_force(1, f)
    assert abs(fix.cv("LQ6.AVE") - lq.mean()) <= CV_TOL * lq.mean()
    lq_gpu = fix.extract("LQ6.local_q", n)
    assert np.abs(lq_gpu - lq).max() <= CV_TOL * lq.max()
    w = fix.cv_weight("LQ6.AVE")
    assert w != 0.0
    F_ref = -env.dLQ_dx(6, w)
    assert rel_force_err(f, F_ref) <= F_TOL


def test_ld_unbiased_alongside_forces(mdt, orc):
    """An unbiased local Q6 on the biased Q6/Q4 environment (the headline workload's
    arrangement): its value and per-atom lq come out of the force pass, every step,
    through post_force and through the captured graph."""
    s = fcc_system(mdt, 8, sigma=0.3, sigma_above=0.1, seed=9)
    x, off, idx, own = s.arrays()
    n = s.n_local
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    qb, lq = env.ld(6)
    q6, q4 = env.q_ave(6), env.q_ave(4)
    args = ("GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE NONE "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME LQ6 STEINH LD 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE SYMBOL v3 LQ6.AVE DIM 1 0 2 2000 v1+v2 WT 0")
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, args)
    fix.init()
    fix.add_hills(np.array([[q6 + q4 + 0.004]]), np.array([0.005]), 0.01)
    for step in (1, 2):
        f = np.zeros((n, 3))
        fix.post_force(step, f)
        assert abs(fix.cv("LQ6.AVE") - lq.mean()) <= CV_TOL * lq.mean(), step
        assert np.abs(fix.extract("LQ6.local_q", n) - lq).max() <= CV_TOL * lq.max()
        assert abs(fix.cv("Q6.AVE") - q6) <= CV_TOL * q6
        F_ref = -(env.dQ_dx(6, fix.cv_weight("Q6.AVE")) + env.dQ_dx(4, fix.cv_weight("Q4.AVE")))
        assert rel_force_err(f, F_ref) <= F_TOL
        assert fix.cv_weight("LQ6.AVE") == 0.0
    fix.zero_forces()
    fix.run(3, 2)
    fix.sync()
    assert abs(fix.cv("LQ6.AVE") - lq.mean()) <= CV_TOL * lq.mean()


def test_ld_unbiased_moving_atoms_every_step(mdt, orc):
    """The fused local Q6 with atoms moving between steps: the next step's reduce kernel
    sums a step's partials and a state read sums the last step's, so the value read after
    each step (steps launched one by one, no graph) is that step's, against the oracle on
    the step's positions, and the same bits as through post_force."""
    s = fcc_system(mdt, 8, sigma=0.3, sigma_above=0.1, seed=19)
    x, off, idx, own = s.arrays()
    n = s.n_local
    mv = Mover(x, n, own)
    rng = np.random.default_rng(6)
    args = ("GAUSSIAN 0.01 0.005 10 PACE 2 HILLS_FILE NONE "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME LQ6 STEINH LD 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE SYMBOL v3 LQ6.AVE DIM 1 0 2 2000 v1+v2 WT 1")
    ctx_a, ctx_b = _ctx(mdt, s), _ctx(mdt, s)
    fa, fb = mdt.Fix(ctx_a, args), mdt.Fix(ctx_b, args)
    fa.init(); fb.init()
    xl = x[:n].copy()
    for step in range(5):
        xa = mv.set_local(xl)
        ctx_a.update_positions(0, xa)
        ctx_b.update_positions(0, xa)
        fa.run(step, 1, 0)
        fa.sync()
        fb.post_force(step)
        lq = orc.Env(xa, off, idx, own, n, 3.9, 18).ld(6)[1]
        print(f"[ld moving] step={step} LQ6={fa.cv('LQ6.AVE'):.8f} ref={lq.mean():.8f}", flush=True)
        assert abs(fa.cv("LQ6.AVE") - lq.mean()) <= CV_TOL * lq.mean(), step
        assert fa.cv("LQ6.AVE") == fb.cv("LQ6.AVE")
        xl = xl + rng.normal(scale=0.05, size=xl.shape)
    fa.close(); fb.close(); ctx_a.close(); ctx_b.close()


def test_lockstep_grid_c1(mdt, orc):
    """Lock-step replay: moving atoms, hills every 10 steps (PACE 10), grid, V, dV/ds, forces each step.

    The oracle keeps its own grid fed with its own CVs.  Forces are compared with the
    oracle's exact gradient contracted with the GPU's chain weight (1e-4 relative); the
    weight itself is compared with the oracle's dV/ds on its natural scale h/sigma
    (near a hill maximum dV/ds passes through zero, where a relative test is meaningless).
    """
    s = fcc_system(mdt, 10, sigma=0.1, seed=21)
    x, off, idx, own = s.arrays()
    n = s.n_local
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE NONE "
                       "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                       "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0")
    mv = Mover(x, n, own)
    rng = np.random.default_rng(0)
    g = orc.Grid([0.0], [1.0], [1000])
    xl = x[:n].copy()
    worst = 0.0
    for step in range(0, 60):
        xa = mv.set_local(xl)
        ctx.update_positions(0, xa)
        f = np.zeros((n, 3))
        fix.post_force(step, f)
        env = orc.Env(xa, off, idx, own, n, 3.9, 18)
        q6 = env.q_ave(6)
        assert abs(fix.cv("Q6.AVE") - q6) <= CV_TOL * q6
        if step % 10 == 0:
            g.deposit([q6], g.wt_height([q6], 0.005, 10.0, 300.0, 8.617343e-5, 0), 0.01)
        V, dV = g.eval([q6])
        assert abs(fix.scalar() - V) <= 1e-4 * abs(V)
        w = fix.cv_weight("Q6.AVE")
        assert abs(w - dV[0]) <= 1e-4 * 0.005 / 0.01, (step, w, dV[0])
        F_ref = -env.dQ_dx(6, w)
        e = rel_force_err(f, F_ref)
        worst = max(worst, e)
        assert e <= F_TOL, (step, e)
        xl = xl + rng.normal(scale=0.005, size=xl.shape)
    grid = fix.grid(1000)
    assert np.abs(grid - g.g).max() <= 1e-4 * np.abs(g.g).max()
    print(f"lock-step worst force rel err {worst:.3e}")


def test_lockstep_c1_ten_thousand_steps(mdt, orc):
    """The 4,000-atom case for its full 10,000 steps of moving atoms (PACE 10, 1,000 deposits):
    Q6.AVE, V and dV/ds against the oracle every step, forces every 100th step, the grid at
    the end."""
    s = fcc_system(mdt, 10, sigma=0.1, seed=21)
    x, off, idx, own = s.arrays()
    n = s.n_local
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE NONE "
                       "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                       "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0")
    mv = Mover(x, n, own)
    rng = np.random.default_rng(3)
    g = orc.Grid([0.0], [1.0], [1000])
    xl = x[:n].copy()
    worst = cvw = 0.0
    f = np.zeros((n, 3))
    for step in range(10000):
        xa = mv.set_local(xl)
        ctx.update_positions(0, xa)
        f[:] = 0.0
        fix.post_force(step, f)
        env = orc.Env(xa, off, idx, own, n, 3.9, 18)
        q6 = env.q_ave(6)
        cvw = max(cvw, abs(fix.cv("Q6.AVE") - q6) / q6)
        assert abs(fix.cv("Q6.AVE") - q6) <= CV_TOL * q6, step
        if step % 10 == 0:
            g.deposit([q6], g.wt_height([q6], 0.005, 10.0, 300.0, 8.617343e-5, 0), 0.01)
        V, dV = g.eval([q6])
        assert abs(fix.scalar() - V) <= 1e-4 * abs(V), step
        w = fix.cv_weight("Q6.AVE")
        assert abs(w - dV[0]) <= 1e-4 * 0.005 / 0.01, (step, w, dV[0])
        if step % 100 == 0:
            e = rel_force_err(f, -env.dQ_dx(6, w))
            worst = max(worst, e)
            assert e <= F_TOL, (step, e)
        if step % 2000 == 0:
            print(f"[c1 lock-step] step {step}: CV rel err so far {cvw:.3e}, force {worst:.3e}", flush=True)
        xl = xl + rng.normal(scale=0.005, size=xl.shape)
    assert fix.hills(1).shape[0] == 1000
    grid = fix.grid(1000)
    assert np.abs(grid - g.g).max() <= 1e-4 * np.abs(g.g).max()
    print(f"[c1 lock-step] 10000 steps: worst CV rel err {cvw:.3e}, worst force rel err {worst:.3e}, "
          f"grid max abs err {np.abs(grid - g.g).max():.3e} of {np.abs(g.g).max():.3e}", flush=True)


def test_graph_run_matches_post_force(mdt, orc):
    s = fcc_system(mdt, 8, sigma=0.2, seed=2)
    n = s.n_local
    res = []
    for mode in ("post", "graph"):
        ctx = _ctx(mdt, s)
        fix = mdt.Fix(ctx, q46_fix(pace=3, wt=1))
        fix.zero_forces()
        if mode == "post":
            for st in range(12):
                fix.post_force(st)
        else:
            fix.run(0, 12)
            fix.sync()
        res.append((fix.forces(n), fix.scalar(), fix.cv("Q6.AVE"), fix.hills(1)))
        fix.close()
        ctx.close()
    (f0, v0, q0, h0), (f1, v1, q1, h1) = res
    assert np.array_equal(f0, f1)
    assert v0 == v1 and q0 == q1
    assert np.array_equal(h0, h1) and h0.shape[0] == 4


def test_hills_file_and_restart(mdt, orc, tmp_path):
    s = fcc_system(mdt, 6, sigma=0.2, seed=3)
    n = s.n_local
    path = str(tmp_path / "HILLS")
    args = (f"GAUSSIAN 0.01 0.005 10 PACE 2 HILLS_FILE {path} "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CV_dim 1 SIMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 1 TEMP 300")
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, args + " METAD_RESTART 0")
    for st in range(10):
        fix.post_force(st)
    g1 = fix.grid(1000)
    fix.close()
    lines = [l for l in open(path) if not l.startswith("#")]
    assert len(lines) == 5
    first = lines[0].split()
    assert first[0] == "0" and abs(float(first[2]) - 0.005) < 1e-15 and float(first[3]) == 0.01
    # restart replays the file into a fresh grid
    ctx2 = _ctx(mdt, s)
    fix2 = mdt.Fix(ctx2, args + " METAD_RESTART 1")
    fix2.init()
    g2 = fix2.grid(1000)
    assert np.abs(g1 - g2).max() <= 1e-12 * np.abs(g1).max()


@pytest.mark.parametrize("expr", ["v1*v2+0.5*exp(v1)", "sqrt(v1^2+v2^2)", "2*v1-v2/3+0.25"])
def test_dim_expression_and_chain_rule(mdt, orc, expr):
    """DIM expressions (interpreted or affine) and exact ds/dv in the chain rule (DimSet.cu:83-97)."""
    s = fcc_system(mdt, 6, sigma=0.2, seed=6)
    x, off, idx, own = s.arrays()
    n = s.n_local
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    v1, v2 = env.q_ave(6), env.q_ave(4)
    f = {"v1*v2+0.5*exp(v1)": (lambda a, b: a * b + 0.5 * np.exp(a), lambda a, b: (b + 0.5 * np.exp(a), a)),
         "sqrt(v1^2+v2^2)": (lambda a, b: np.sqrt(a * a + b * b),
                             lambda a, b: (a / np.sqrt(a * a + b * b), b / np.sqrt(a * a + b * b))),
         "2*v1-v2/3+0.25": (lambda a, b: 2 * a - b / 3 + 0.25, lambda a, b: (2.0, -1.0 / 3.0))}[expr]
    s0 = f[0](v1, v2)
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, q46_fix(expr=f'"{expr}"', lo=0.0, hi=3.0, nb=3000))
    fix.init()
    fix.add_hills(np.array([[s0 + 0.007]]), np.array([0.005]), 0.01)
    fix.post_force(1)
    assert abs(fix.vector(0) - s0) <= 1e-5 * abs(s0)
    V, dV = fix.bias_eval(np.array([[fix.vector(0)]]))
    d1, d2 = f[1](v1, v2)
    assert abs(fix.cv_weight("Q6.AVE") - dV[0, 0] * d1) <= 1e-4 * abs(dV[0, 0] * d1)
    assert abs(fix.cv_weight("Q4.AVE") - dV[0, 0] * d2) <= 1e-4 * abs(dV[0, 0] * d2)
