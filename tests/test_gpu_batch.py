"""GPU parity of the structure-factor CV and of batched walkers (run with -m gpu)."""
import numpy as np
import pytest

from helpers import Mover, fcc_system, liquid_system, rel_force_err

pytestmark = pytest.mark.gpu

CV_TOL = 1e-5
F_TOL = 1e-4


def _kvecs(seed, nrand, a=5.26):
    k0 = 2 * np.pi / a
    ks = [(sx * k0, sy * k0, sz * k0) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    rng = np.random.default_rng(seed)
    for _ in range(nrand):
        m = rng.uniform(1.5, 3.5)
        v = rng.normal(size=3)
        ks.append(tuple(m * v / np.linalg.norm(v)))
    return np.array(ks)


@pytest.mark.parametrize("nrand,kscale", [(40, 1.0), (300, 1.0), (300, 6.0)])
def test_structure_factor_value_and_forces(mdt, orc, nrand, kscale):
    """48 k vectors take the FP64 phase; 308 (above 256) the tile-relative FP32 phase
    over atoms in Hilbert order (the force written through that order), with the
    phases used as they are, or, with |k| six times larger (tile phases beyond 128
    revolutions), reduced to [0, 1) first."""
    s = fcc_system(mdt, 6, sigma=0.3, seed=12)
    x, off, idx, own = s.arrays()
    n = s.n_local
    k = _kvecs(3, nrand) * kscale
    kv = " ".join(f"KVEC {a:.17g} {b:.17g} {c:.17g}" for a, b, c in k)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.5 0.005 10 PACE 10 HILLS_FILE NONE CAL NAME SK SK all {kv} "
                       "CV_dim 1 SYMBOL v1 SK.AVE DIM 1 0 400 4000 v1 WT 0")
    v_ref, A, S = orc.sk(x[:n], k)
    print(f"[sk] n={n} nk={len(k)} S.AVE={v_ref:.4f}", flush=True)
    fix.init()
    fix.add_hills(np.array([[v_ref + 0.3]]), np.array([0.005]), 0.5)
    f = np.zeros((n, 3))
    fix.post_force(1, f)
    assert abs(fix.cv("SK.AVE") - v_ref) <= CV_TOL * abs(v_ref)
    w = fix.cv_weight("SK.AVE")
    assert w != 0.0
    F_ref = -orc.sk_grad(x[:n], k, w)
    assert rel_force_err(f, F_ref) <= F_TOL


@pytest.mark.parametrize("nrand", [56, 300])
def test_structure_factor_large_phase(mdt, orc, nrand):
    """Phases of ~1e3 rad (500k-atom slab scale): the FP64 phase reduction (64 k) and the
    tile-relative phases (308 k) keep S(k) at 1e-5 and the forces at 1e-4."""
    L = 273.52
    s = mdt.System().set_box([0, 0, 0], [L, L, L]).add_uniform(20000, 0.0, L, seed=4).build(5.9)
    x = s.arrays()[0][: s.n_local]
    n = s.n_local
    k = _kvecs(9, nrand)
    kv = " ".join(f"KVEC {a:.17g} {b:.17g} {c:.17g}" for a, b, c in k)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.5 0.005 10 PACE 0 HILLS_FILE NONE CAL NAME SK SK all {kv} "
                       "CV_dim 1 SYMBOL v1 SK.AVE DIM 1 0 50 1000 v1 WT 0")
    v_ref = orc.sk(x, k)[0]
    print(f"[sk large phase] n={n} nk={len(k)} S.AVE={v_ref:.4f}", flush=True)
    fix.init()
    fix.add_hills(np.array([[v_ref + 0.3]]), np.array([0.005]), 0.5)
    f = np.zeros((n, 3))
    fix.post_force(0, f)
    assert abs(fix.cv("SK.AVE") - v_ref) <= CV_TOL * abs(v_ref)
    w = fix.cv_weight("SK.AVE")
    assert w != 0.0
    assert rel_force_err(f, -orc.sk_grad(x, k, w)) <= F_TOL


def _walker_systems(mdt):
    return [fcc_system(mdt, 8, sigma=0.1, seed=31),          # 2,048 atoms FCC
            liquid_system(mdt, 2500, density=0.021, seed=32),
            fcc_system(mdt, 10, sigma=0.1, seed=33),         # 4,000 atoms FCC
            liquid_system(mdt, 3100, density=0.021, seed=34)]


def test_batched_walkers_2d_grids(mdt, orc):
    systems = _walker_systems(mdt)
    nw = len(systems)
    ctx = mdt.Context(0, nw)
    for w, s in enumerate(systems):
        ctx.bind_system(w, s)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 2 HILLS_FILE NONE TEMP 300 "
                       "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
                       "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                       "CV_dim 2 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE "
                       "DIM 1 0 1 200 v1 DIM 2 0 1 200 v2 WT 1")
    ntot = sum(s.n_local for s in systems)
    arrs = [s.arrays() for s in systems]
    envs = [orc.Env(x, off, idx, own, s.n_local, 3.9, 18) for s, (x, off, idx, own) in zip(systems, arrs)]
    grids = [orc.Grid([0.0, 0.0], [1.0, 1.0], [200, 200]) for _ in systems]
    for step in range(7):
        f = np.zeros((ntot, 3))
        fix.post_force(step, f)
        o = 0
        for w, (s, env, g) in enumerate(zip(systems, envs, grids)):
            n = s.n_local
            q4, q6 = env.q_ave(4), env.q_ave(6)
            assert abs(fix.cv("Q4.AVE", w) - q4) <= CV_TOL * q4
            assert abs(fix.cv("Q6.AVE", w) - q6) <= CV_TOL * q6
            sv = [q4, q6]
            if step % 2 == 0:
                g.deposit(sv, g.wt_height(sv, 0.005, 10.0, 300.0, 8.617343e-5, 1), 0.01)
            V, dV = g.eval(sv)
            assert abs(fix.scalar(w) - V) <= 1e-4 * abs(V)
            w4, w6 = fix.cv_weight("Q4.AVE", w), fix.cv_weight("Q6.AVE", w)
            # chain weights on the natural dV/ds scale h/sigma (see test_lockstep_grid_c1)
            assert abs(w4 - dV[0]) <= 1e-4 * 0.005 / 0.01 and abs(w6 - dV[1]) <= 1e-4 * 0.005 / 0.01
            F_ref = -(env.dQ_dx(4, w4) + env.dQ_dx(6, w6))
            assert rel_force_err(f[o:o + n], F_ref) <= F_TOL, (w, step)
            o += n
    for w, g in enumerate(grids):
        gg = fix.grid(200 * 200, w)
        assert np.abs(gg - g.g).max() <= 1e-4 * np.abs(g.g).max()


def test_walkers_graph_equals_single(mdt, orc):
    """The batched graph gives every walker exactly what a one-walker context gives it."""
    systems = _walker_systems(mdt)[:3]
    args = ("GAUSSIAN 0.01 0.005 10 PACE 3 HILLS_FILE NONE "
            "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CV_dim 2 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE DIM 1 0 1 200 v1 DIM 2 0 1 200 v2 WT 1")
    ctx = mdt.Context(0, len(systems))
    for w, s in enumerate(systems):
        ctx.bind_system(w, s)
    fix = mdt.Fix(ctx, args)
    fix.zero_forces()
    fix.run(0, 9)
    fix.sync()
    for w, s in enumerate(systems):
        c1 = mdt.Context(0, 1)
        c1.bind_system(0, s)
        f1 = mdt.Fix(c1, args)
        f1.zero_forces()
        f1.run(0, 9)
        f1.sync()
        assert np.array_equal(f1.forces(s.n_local), fix.forces(s.n_local, w))
        assert f1.scalar() == fix.scalar(w)
        f1.close()
        c1.close()


def test_run_batches_equal_single_launches(mdt):
    """Steps replayed as graphs of 8 captured steps, and steps whose kernels are launched
    one by one, give the same bits as one graph launch per step: forces, CVs, the bias
    state and the deposited hills (well-tempered, PACE 3, 23 steps: batches, a remainder,
    and deposits inside and across batches)."""
    s = fcc_system(mdt, 6, sigma=0.15, seed=52)
    n = s.n_local
    args = ("GAUSSIAN 0.01 0.005 10 PACE 3 TEMP 300 HILLS_FILE NONE "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE DIM 1 0 2 2000 v1+v2 WT 1")
    out = []
    for per_launch in (1, 8, None, 0):
        ctx = mdt.Context(0, 1)
        ctx.bind_system(0, s)
        fix = mdt.Fix(ctx, args)
        fix.init()
        fix.zero_forces()
        fix.run(0, 23, per_launch)
        fix.sync()
        out.append((fix.forces(n), fix.cv("Q6.AVE"), fix.cv("Q4.AVE"), fix.scalar(), fix.hills(1)))
        print(f"[run batches] per_launch={per_launch} V={out[-1][3]:.6g} hills={len(out[-1][4])}", flush=True)
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0])
        assert o[1] == out[0][1] and o[2] == out[0][2] and o[3] == out[0][3]
        assert np.array_equal(o[4], out[0][4])
    assert len(out[0][4]) == 8          # steps 0, 3, ..., 21
