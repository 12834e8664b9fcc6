"""The reference's sparse-hash bias (Gaussian_Hill_type 1, src/grids/GH_t1_sparseHash.cu) as a
dense 3-D device grid, against the oracle's recursive restatement (run with -m gpu).

Bars: grid values and V within 1e-12 of the FP64 oracle relative to their maximum (the same
arithmetic, contraction off; only exp may differ in the last bit); dV/ds within 1e-10; CVs
1e-5 and forces 1e-4 relative as in test_gpu_parity.py.
"""
import numpy as np
import pytest

from helpers import fcc_system, rel_force_err

pytestmark = pytest.mark.gpu

CV_TOL = 1e-5
F_TOL = 1e-4
KT = 300.0 * 8.617343e-5

LO, HI, NB = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [50, 40, 30]
FIX = ("GAUSSIAN 0.03 0.005 10 PACE {pace} HILLS_FILE NONE TEMP 300 Gaussian_Hill_type 1 "
       "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
       "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
       "CAL NAME Q3 STEINH Q 3 all cutoff_r 3.9 cutoff_Natoms 18 "
       "CV_dim 3 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE SYMBOL v3 Q3.AVE "
       "DIM 1 0 1 50 v1 DIM 2 0 1 40 v2 DIM 3 0 1 30 v3 WT {wt}")


def _fix(mdt, system, pace, wt):
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, system)
    return ctx, mdt.Fix(ctx, FIX.format(pace=pace, wt=wt))


def test_sparse_hash_grid_deposit_and_eval(mdt, orc):
    """Hills deposited through the C ABI land on the same bins with the same values as the
    oracle's; V and dV/ds at points inside, at the clamped edges and outside agree."""
    s = fcc_system(mdt, 4, sigma=0.1, seed=61)
    ctx, fix = _fix(mdt, s, 0, 0)
    fix.init()
    rng = np.random.default_rng(62)
    c = rng.uniform(0.05, 0.95, size=(25, 3))
    h = rng.uniform(0.001, 0.006, size=25)
    print(f"[grid3d] deposit 25 hills on {NB}", flush=True)
    fix.add_hills(c, h, 0.03)
    g = orc.GridT1(LO, HI, NB)
    for k in range(25):
        g.deposit(c[k], h[k], 0.03)
    gg = fix.grid(int(np.prod(NB)))
    assert np.abs(gg - g.g).max() <= 1e-12 * np.abs(g.g).max()
    pts = np.vstack([rng.uniform(0.0, 1.0, size=(150, 3)), rng.uniform(-0.2, 1.2, size=(50, 3))])
    V, dV = fix.bias_eval(pts)
    ref = [g.eval(p) for p in pts]
    Vx = np.array([r[0] for r in ref])
    dVx = np.array([r[1] for r in ref])
    assert np.abs(V - Vx).max() <= 1e-12 * np.abs(Vx).max()
    assert np.abs(dV - dVx).max() <= 1e-10 * np.abs(dVx).max()
    fix.close()
    ctx.close()


def test_sparse_hash_grid_well_tempered_steps_and_forces(mdt, orc):
    """Five steps with a deposit on each (PACE 1, well-tempered): the heights follow
    h0 exp(-V/(kT (gamma-1))) with V the 3-D spline before the deposit, the grid matches the
    oracle's replay, and the forces are the chain rule of the final dV/ds through Q4, Q6, Q3."""
    s = fcc_system(mdt, 5, sigma=0.12, seed=63)
    x, off, idx, own = s.arrays()
    n = s.n_local
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    q = np.array([env.q_ave(4), env.q_ave(6), env.q_ave(3)])
    ctx, fix = _fix(mdt, s, 1, 1)
    fix.init()
    f = np.zeros((n, 3))
    for step in range(5):
        f[:] = 0.0
        print(f"[grid3d] step {step}", flush=True)
        fix.post_force(step, f)
    sv = np.array([fix.vector(d) for d in range(3)])
    assert np.abs(sv - q).max() <= CV_TOL * np.abs(q).max()
    g = orc.GridT1(LO, HI, NB)
    heights = []
    for _ in range(5):
        w = g.wt_height(sv, 0.005, 10.0, 300.0, 8.617343e-5, 1)
        heights.append(w)
        g.deposit(sv, w, 0.03)
    hills = fix.hills(3)
    assert hills.shape[0] == 5
    assert np.abs(hills[:, 4] - np.array(heights)).max() <= 1e-12 * 0.005
    assert heights[-1] < heights[0]
    gg = fix.grid(int(np.prod(NB)))
    assert np.abs(gg - g.g).max() <= 1e-12 * np.abs(g.g).max()
    V, dVds = g.eval(sv)
    w4, w6, w3 = fix.cv_weight("Q4.AVE"), fix.cv_weight("Q6.AVE"), fix.cv_weight("Q3.AVE")
    assert np.abs(np.array([w4, w6, w3]) - dVds).max() <= 1e-8 * np.abs(dVds).max()
    F_ref = -(env.dQ_dx(4, w4) + env.dQ_dx(6, w6) + env.dQ_dx(3, w3))
    assert rel_force_err(f, F_ref) <= F_TOL
    fix.close()
    ctx.close()


def test_sparse_hash_grid_restart_replays_hills(mdt, orc, tmp_path):
    """A 3-D HILLS file written by one run rebuilds the same grid with METAD_RESTART 1."""
    s = fcc_system(mdt, 4, sigma=0.1, seed=64)
    path = str(tmp_path / "HILLS3")
    args = FIX.format(pace=1, wt=1).replace("HILLS_FILE NONE", f"HILLS_FILE {path}")
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, args)
    fix.init()
    for step in range(4):
        fix.post_force(step)
    g1 = fix.grid(int(np.prod(NB)))
    fix.close()
    rows = np.loadtxt(path, comments="#")
    assert rows.shape == (4, 6)            # step, s1, s2, s3, height, sigma
    ctx2 = mdt.Context(0, 1)
    ctx2.bind_system(0, s)
    fix2 = mdt.Fix(ctx2, args + " METAD_RESTART 1")
    fix2.init()
    g2 = fix2.grid(int(np.prod(NB)))
    assert np.abs(g1 - g2).max() <= 1e-12 * np.abs(g1).max()
    fix2.close()
    ctx.close()
    ctx2.close()
