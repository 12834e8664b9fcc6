"""GPU parity at the edges: group-restricted CALs, isolated atoms, long lists with a large
cap, odd degree (Q3) forces, and walkers from empty to 4,000 atoms (run with -m gpu).

Bars as in test_gpu_parity.py: selection bit-exact, CV 1e-5 relative, forces 1e-4
relative (max |F - F_ref| / max |F_ref|).
"""
import numpy as np
import pytest

from helpers import fcc_system, liquid_system, rel_force_err

pytestmark = pytest.mark.gpu

CV_TOL = 1e-5
F_TOL = 1e-4


def _one(mdt, system, args, group=None):
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, system)
    if group is not None:
        ctx.bind_group(0, group)
    return ctx, mdt.Fix(ctx, args)


def _biased_step(fix, s_center, n):
    """One step with a hill placed next to the CV value, so dV/ds is well away from zero."""
    fix.init()
    fix.add_hills(np.array([[s_center]]), np.array([0.005]), 0.01)
    f = np.zeros((n, 3))
    fix.post_force(1, f)
    return f


def test_group_restricted_cal(mdt, orc):
    """STEINH on a group: centers are the group's atoms, Q6.AVE divides by the group size,
    neighbours are any atoms, and forces act on group atoms only (the reference applies
    the bias to group atoms, CVs_Steinhardt.cu:841-868)."""
    s = fcc_system(mdt, 8, sigma=0.1, seed=41)
    x, off, idx, own = s.arrays()
    n = s.n_local
    grp = (x[:n, 2] < x[:n, 2].mean()).astype(np.uint8)
    print(f"[group] n={n} in group={int(grp.sum())}", flush=True)
    env = orc.Env(x, off, idx, own, n, 3.9, 18, group=grp)
    q_ref = env.q_ave(6)
    ctx, fix = _one(mdt, s, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                            "CAL NAME Q6 STEINH Q 6 lower cutoff_r 3.9 cutoff_Natoms 18 "
                            "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0", group=grp)
    f = _biased_step(fix, q_ref + 0.004, n)
    assert abs(fix.cv("Q6.AVE") - q_ref) <= CV_TOL * q_ref
    w = fix.cv_weight("Q6.AVE")
    assert w != 0.0
    F_ref = -env.dQ_dx(6, w)
    inside = grp.astype(bool)
    assert rel_force_err(f[inside], F_ref[inside]) <= F_TOL
    assert np.all(f[~inside] == 0.0)


def test_isolated_atoms(mdt, orc):
    """A dilute gas: many atoms have no neighbour within rc (q_l = 0, no force), and count
    in the average as the reference counts them."""
    s = liquid_system(mdt, 3000, density=0.004, seed=42)
    x, off, idx, own = s.arrays()
    n = s.n_local
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    assert (env.nsel == 0).sum() > n // 4
    q_ref = env.q_ave(6)
    print(f"[isolated] n={n} empty={(env.nsel == 0).sum()} Q6={q_ref:.6f}", flush=True)
    ctx, fix = _one(mdt, s, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                            "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0")
    f = _biased_step(fix, q_ref + 0.004, n)
    assert abs(fix.cv("Q6.AVE") - q_ref) <= CV_TOL * q_ref
    assert np.array_equal(fix.extract("Q6.nsel", n).astype(np.int32), env.nsel)
    w = fix.cv_weight("Q6.AVE")
    F_ref = -env.dQ_dx(6, w)
    assert rel_force_err(f, F_ref) <= F_TOL
    assert np.all(f[env.nsel + np.bincount(env.owner_of(env.sel_idx[env.sel_idx >= 0]), minlength=n) == 0] == 0.0)


def test_long_lists_large_cap(mdt, orc):
    """Lists longer than 64 entries (both halves of the 128-bit masks) and a cap of 40
    within 5 A that binds for many atoms: selection bit-exact, CVs and forces."""
    s = liquid_system(mdt, 4000, density=0.09, seed=43, rg=6.2)
    x, off, idx, own = s.arrays()
    n = s.n_local
    c = np.diff(off)
    N, rc = 40, 5.0
    env = orc.Env(x, off, idx, own, n, rc, N)
    print(f"[long] n={n} list max={c.max()} mean={c.mean():.1f} capped={(env.cnt > N).sum()}", flush=True)
    assert c.max() > 64 and c.max() <= 128
    assert (env.cnt > N).sum() > 50
    args = ("GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
            f"CAL NAME Q6 STEINH Q 6 all cutoff_r {rc} cutoff_Natoms {N} "
            f"CAL NAME Q4 STEINH Q 4 all cutoff_r {rc} cutoff_Natoms {N} "
            "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE DIM 1 0 2 2000 v1+v2 WT 0")
    ctx, fix = _one(mdt, s, args)
    fix.init()
    fix.post_force(0)
    sel, cnt, ns = fix.selection("Q6", n, N)
    assert np.array_equal(sel, env.sel_idx)
    assert np.array_equal(fix.extract("Q6.nsel", n).astype(np.int32), env.nsel)
    q6, q4 = env.q_ave(6), env.q_ave(4)
    f = np.zeros((n, 3))
    fix.add_hills(np.array([[q6 + q4 + 0.004]]), np.array([0.005]), 0.01)
    fix.post_force(1, f)
    assert abs(fix.cv("Q6.AVE") - q6) <= CV_TOL * q6
    assert abs(fix.cv("Q4.AVE") - q4) <= CV_TOL * q4
    w6, w4 = fix.cv_weight("Q6.AVE"), fix.cv_weight("Q4.AVE")
    F_ref = -(env.dQ_dx(6, w6) + env.dQ_dx(4, w4))
    assert rel_force_err(f, F_ref) <= F_TOL


def test_odd_degree_forces(mdt, orc):
    """Q3 (odd l: the reverse pair enters with -1) together with Q6 on one environment."""
    s = fcc_system(mdt, 6, sigma=0.15, seed=44)
    x, off, idx, own = s.arrays()
    n = s.n_local
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    q3, q6 = env.q_ave(3), env.q_ave(6)
    ctx, fix = _one(mdt, s, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                            "CAL NAME Q3 STEINH Q 3 all cutoff_r 3.9 cutoff_Natoms 18 "
                            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                            "CV_dim 1 SYMBOL v1 Q3.AVE SYMBOL v2 Q6.AVE DIM 1 0 2 2000 v1+v2 WT 0")
    f = _biased_step(fix, q3 + q6 + 0.004, n)
    assert abs(fix.cv("Q3.AVE") - q3) <= CV_TOL * q3
    w3, w6 = fix.cv_weight("Q3.AVE"), fix.cv_weight("Q6.AVE")
    assert w3 != 0.0
    F_ref = -(env.dQ_dx(3, w3) + env.dQ_dx(6, w6))
    assert rel_force_err(f, F_ref) <= F_TOL


def test_walkers_from_tiny_to_large(mdt, orc):
    """Walkers of 108, 4,000 and 500 atoms in one context: each matches the oracle on its own."""
    systems = [fcc_system(mdt, 3, sigma=0.1, seed=45), fcc_system(mdt, 10, sigma=0.1, seed=46),
               fcc_system(mdt, 5, sigma=0.12, seed=47)]
    ctx = mdt.Context(0, len(systems))
    for w, s in enumerate(systems):
        ctx.bind_system(w, s)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                       "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                       "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0")
    fix.init()
    envs = []
    for w, s in enumerate(systems):
        x, off, idx, own = s.arrays()
        env = orc.Env(x, off, idx, own, s.n_local, 3.9, 18)
        envs.append(env)
        fix.add_hills(np.array([[env.q_ave(6) + 0.004]]), np.array([0.005]), 0.01, walker=w)
    ntot = sum(s.n_local for s in systems)
    f = np.zeros((ntot, 3))
    fix.post_force(1, f)
    o = 0
    for w, (s, env) in enumerate(zip(systems, envs)):
        n = s.n_local
        q = env.q_ave(6)
        assert abs(fix.cv("Q6.AVE", w) - q) <= CV_TOL * q, w
        F_ref = -env.dQ_dx(6, fix.cv_weight("Q6.AVE", w))
        assert rel_force_err(f[o:o + n], F_ref) <= F_TOL, w
        o += n


@pytest.mark.parametrize("q4_group", ["upper", "all"])
def test_cals_on_different_groups(mdt, orc, q4_group):
    """Q6 on group 'lower' and Q4 on another group (its complement, or all), both biased
    through v1+v2: each CV averages over its own group (its own count) and its force acts on
    its own group's atoms, as the reference pools one environment per group
    (CVs_Steinhardt.cu:191-220) and applies each CV's bias to its group (:841-868)."""
    s = fcc_system(mdt, 8, sigma=0.15, seed=43)
    x, off, idx, own = s.arrays()
    n = s.n_local
    lower = (x[:n, 2] < x[:n, 2].mean()).astype(np.uint8)
    upper = (1 - lower).astype(np.uint8)
    g4 = upper if q4_group == "upper" else None
    env6 = orc.Env(x, off, idx, own, n, 3.9, 18, group=lower)
    env4 = orc.Env(x, off, idx, own, n, 3.9, 18, group=g4)
    q6, q4 = env6.q_ave(6), env4.q_ave(4)
    print(f"[groups] n={n} lower={int(lower.sum())} Q6(lower)={q6:.6f} Q4({q4_group})={q4:.6f}", flush=True)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    ctx.bind_group_named(0, "lower", lower)
    ctx.bind_group_named(0, "upper", upper)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                       "CAL NAME Q6 STEINH Q 6 lower cutoff_r 3.9 cutoff_Natoms 18 "
                       f"CAL NAME Q4 STEINH Q 4 {q4_group} cutoff_r 3.9 cutoff_Natoms 18 "
                       "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE DIM 1 0 2 2000 v1+v2 WT 0")
    f = _biased_step(fix, q6 + q4 + 0.004, n)
    assert abs(fix.cv("Q6.AVE") - q6) <= CV_TOL * q6
    assert abs(fix.cv("Q4.AVE") - q4) <= CV_TOL * q4
    w = fix.cv_weight("Q6.AVE")
    assert w != 0.0 and w == fix.cv_weight("Q4.AVE")
    m6 = lower.astype(bool)[:, None]
    m4 = np.ones((n, 1), bool) if g4 is None else upper.astype(bool)[:, None]
    F_ref = -(np.where(m6, env6.dQ_dx(6, w), 0.0) + np.where(m4, env4.dQ_dx(4, w), 0.0))
    assert rel_force_err(f, F_ref) <= F_TOL
    fix.close()
    ctx.close()


def test_two_biased_cals_with_one_definition(mdt, orc):
    """Two CALs with identical STEINH Q 6 definitions, both biased (DIM v1+2*v2): one shared
    environment degree whose force pass takes the sum of the two weights."""
    s = fcc_system(mdt, 7, sigma=0.2, seed=44)
    x, off, idx, own = s.arrays()
    n = s.n_local
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    q6 = env.q_ave(6)
    ctx, fix = _one(mdt, s, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                            "CAL NAME A STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                            "CAL NAME B STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                            "CV_dim 1 SYMBOL v1 A.AVE SYMBOL v2 B.AVE DIM 1 0 3 3000 v1+2*v2 WT 0")
    f = _biased_step(fix, 3 * q6 + 0.006, n)
    assert fix.cv("A.AVE") == fix.cv("B.AVE")
    assert abs(fix.cv("A.AVE") - q6) <= CV_TOL * q6
    wa, wb = fix.cv_weight("A.AVE"), fix.cv_weight("B.AVE")
    assert wa != 0.0 and abs(wb - 2.0 * wa) <= 1e-12 * abs(wb)
    assert rel_force_err(f, -env.dQ_dx(6, wa + wb)) <= F_TOL
    fix.close()
    ctx.close()


def test_reference_q4_deck_long_lists_and_rebuilds(mdt, orc):
    """The reference's own Q4 deck geometry (test/STEINH/Q4/in.test-steinhardt: FCC a = 3.26,
    STEINH Q 4 cutoff_r 6.0 cutoff_Natoms 12, list cutoff 6.0 + 2.0 skin): ~240 list entries
    per atom, beyond the 128 a selection mask holds, so the Steinhardt kernels walk near
    lists cut just above rc (here rc (1 + 1/64): the 48-atom shell at 6.1 A sits right above
    rc).  Selection (slot order), Q4.AVE and forces against the oracle on the bound lists,
    while the atoms move far past the rebuild threshold (half the margin above rc), so the
    near lists are rebuilt inside the steps."""
    from helpers import Mover
    s = fcc_system(mdt, 8, a=3.26, sigma=0.05, seed=48, rg=8.0)
    x, off, idx, own = s.arrays()
    n = s.n_local
    c = np.diff(off)
    print(f"[q4 deck] n={n} list min={c.min()} mean={c.mean():.1f} max={c.max()}", flush=True)
    assert c.min() > 200 and c.max() > 240
    N, rc = 12, 6.0
    ctx, fix = _one(mdt, s, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                            f"CAL NAME Q4 STEINH Q 4 all cutoff_r {rc} cutoff_Natoms {N} "
                            "CV_dim 1 SYMBOL v1 Q4.AVE DIM 1 0 1 1000 v1 WT 0")
    env = orc.Env(x, off, idx, own, n, rc, N)
    fix.init()
    fix.add_hills(np.array([[env.q_ave(4) + 0.004]]), np.array([0.005]), 0.01)
    mv = Mover(x, n, own)
    rng = np.random.default_rng(8)
    xl = x[:n].copy()
    moved = np.zeros(n)
    for step in range(1, 9):
        xa = mv.set_local(xl)
        ctx.update_positions(0, xa)
        f = np.zeros((n, 3))
        fix.post_force(step, f)
        env = orc.Env(xa, off, idx, own, n, rc, N)
        q4 = env.q_ave(4)
        assert abs(fix.cv("Q4.AVE") - q4) <= CV_TOL * q4, step
        assert np.array_equal(fix.extract("Q4.count_rc", n).astype(np.int32), env.cnt), step
        sel, cnt, ns = fix.selection("Q4", n, N)
        assert np.array_equal(sel, env.sel_idx), step
        w = fix.cv_weight("Q4.AVE")
        assert rel_force_err(f, -env.dQ_dx(4, w)) <= F_TOL, step
        d = rng.normal(scale=0.06, size=xl.shape)
        xl = xl + d
        moved = np.linalg.norm(xl - x[:n], axis=1)
    # the run went far past the rebuild threshold (under 0.05 A here) and stayed inside the
    # list skin (1 A)
    assert moved.max() > 0.3 and moved.max() < 1.0, moved.max()


def test_reference_q4_deck_full_size(mdt, orc):
    """The reference's Q4 deck at its own size: 32,000 atoms (20^3 FCC cells, a = 3.26 A,
    jittered 0.02 A), STEINH Q 4 N = 12 rc = 6.0 on 8.0 A lists, PACE 1 with WT: two steps
    (a deposit each) with Q4.AVE, V, dV/ds, the selection in slot order and the forces
    against the oracle."""
    s = mdt.System().add_fcc(20, 20, 20, 3.26, 0.02, 0.02, 0.0, seed=10).build(8.0)
    x, off, idx, own = s.arrays()
    n = s.n_local
    print(f"[q4 deck 32k] n={n} mean list {off[-1] / n:.1f}", flush=True)
    env = orc.Env(x, off, idx, own, n, 6.0, 12)
    q4 = env.q_ave(4)
    ctx, fix = _one(mdt, s, "GAUSSIAN 0.2 0.01 10.0 PACE 1 TEMP 10 HILLS_FILE NONE "
                            "CAL NAME Q4 STEINH Q 4 all cutoff_Natoms 12 cutoff_r 6.0 "
                            "CV_dim 1 SYMBOL v1 Q4.AVE DIM 1 0 40 400 v1 WT 1")
    fix.init()
    g = orc.Grid([0.0], [40.0], [400])
    for step in range(2):
        f = np.zeros((n, 3))
        fix.post_force(step, f)
        assert abs(fix.cv("Q4.AVE") - q4) <= CV_TOL * q4, step
        g.deposit([q4], g.wt_height([q4], 0.01, 10.0, 10.0, 8.617343e-5, 1), 0.2)
        V, dV = g.eval([q4])
        assert abs(fix.scalar() - V) <= 1e-4 * abs(V), step
        w = fix.cv_weight("Q4.AVE")
        assert abs(w - dV[0]) <= 1e-4 * 0.01 / 0.2, (step, w, dV[0])
        assert rel_force_err(f, -env.dQ_dx(4, w)) <= F_TOL, step
    sel, cnt, ns = fix.selection("Q4", n, 12)
    assert np.array_equal(sel, env.sel_idx)
    assert np.array_equal(cnt, env.cnt)
    fix.close()
    ctx.close()


def test_more_than_128_within_the_cutoff_is_refused(mdt):
    """A Steinhardt cutoff with more than 128 list entries inside it (FCC a = 3.26 A, rc 7.5 A:
    ~190 atoms) cannot be held by the selection masks: the fix refuses it with a message at
    init instead of selecting from a truncated list."""
    s = fcc_system(mdt, 8, a=3.26, sigma=0.05, seed=49, rg=8.0)
    ctx, fix = _one(mdt, s, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                            "CAL NAME Q4 STEINH Q 4 all cutoff_r 7.5 cutoff_Natoms 12 "
                            "CV_dim 1 SYMBOL v1 Q4.AVE DIM 1 0 1 1000 v1 WT 0")
    with pytest.raises(mdt.MdtError, match="more than 128 list entries within the Steinhardt cutoff"):
        fix.init()
    fix.close()
    ctx.close()


def test_walkers_with_long_and_short_lists(mdt, orc):
    """Two walkers in one context, one with the Q4 deck's ~240-entry lists (walked through
    near lists) and one with ordinary 5.9 A lists: each walker's Q4.AVE, chain weight and
    forces against the oracle on its own bound lists."""
    s_long = fcc_system(mdt, 8, a=3.26, sigma=0.05, seed=61, rg=8.0)
    s_short = fcc_system(mdt, 6, a=5.26, sigma=0.15, seed=62, rg=6.5)
    systems = [s_long, s_short]
    ctx = mdt.Context(0, 2)
    for w, s in enumerate(systems):
        ctx.bind_system(w, s)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
                       "CAL NAME Q4 STEINH Q 4 all cutoff_r 6.0 cutoff_Natoms 12 "
                       "CV_dim 1 SYMBOL v1 Q4.AVE DIM 1 0 1 1000 v1 WT 0")
    fix.init()
    refs = []
    for w, s in enumerate(systems):
        x, off, idx, own = s.arrays()
        env = orc.Env(x, off, idx, own, s.n_local, 6.0, 12)
        q4 = env.q_ave(4)
        fix.add_hills(np.array([[q4 + 0.004]]), np.array([0.005]), 0.01, walker=w)
        refs.append((env, q4, np.diff(off).max()))
    print(f"[walkers long/short] longest lists {[r[2] for r in refs]}", flush=True)
    assert refs[0][2] > 128
    ns = [s.n_local for s in systems]
    f = np.zeros((sum(ns), 3))
    fix.post_force(1, f)
    a = 0
    for w, (env, q4, _) in enumerate(refs):
        assert abs(fix.cv("Q4.AVE", w) - q4) <= CV_TOL * q4, w
        wt = fix.cv_weight("Q4.AVE", w)
        assert wt != 0.0
        assert rel_force_err(f[a:a + ns[w]], -env.dQ_dx(4, wt)) <= F_TOL, w
        a += ns[w]
    fix.close()
    ctx.close()
