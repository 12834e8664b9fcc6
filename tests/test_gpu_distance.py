"""DISTANCE CV (src/cvs/CVs_DISTANCE.cu) on the GPU against the oracle, alone and biased
together with Q6.AVE on a 2-D grid (run with -m gpu).

Bars: the distance equals the oracle's to 1e-14 relative (the same FP64 operations); forces
1e-4 relative (max |F - F_ref| / max |F_ref|) as in test_gpu_parity.py.
"""
import numpy as np
import pytest

from helpers import fcc_system, rel_force_err

pytestmark = pytest.mark.gpu

F_TOL = 1e-4
A = 5.26


def _nearest(x, p):
    return int(np.argmin(np.sum((x - p) ** 2, axis=1)))


@pytest.mark.parametrize("across", [False, True])
def test_distance_value_and_forces_with_q6(mdt, orc, across):
    s = fcc_system(mdt, 5, sigma=0.1, seed=81)
    x, off, idx, own = s.arrays()
    n = s.n_local
    lo, hi = s.box()
    L = np.asarray(hi) - np.asarray(lo)
    xl = x[:n]
    a1 = _nearest(xl, np.array([0.0, 0.0, 0.0]))
    # across the periodic face: the image of a2 at x - L is the near one
    a2 = _nearest(xl, np.array([L[0] - A / 2, A / 2, 0.0]) if across else np.array([A, A / 2, A / 2]))
    r, d = orc.distance_cv(xl, L, a1, a2)
    raw = xl[a2] - xl[a1]
    assert np.any(np.abs(raw) > L / 2) == across      # the half-box fold is (not) exercised
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    q6 = env.q_ave(6)
    lo_r = round(r, 1) - 1.0
    args = (f"GAUSSIAN 0.01 0.005 10 PACE 0 HILLS_FILE NONE "
            f"CAL NAME D DISTANCE {a1 + 1} {a2 + 1} "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            f"CV_dim 2 SYMBOL v1 D.dist SYMBOL v2 Q6.AVE DIM 1 {lo_r} {lo_r + 2.0} 200 v1 DIM 2 0 1 100 v2 WT 0")
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, args)
    fix.init()
    fix.add_hills(np.array([[r + 0.006, q6 + 0.004]]), np.array([0.005]), 0.01)
    f = np.zeros((n, 3))
    print(f"[distance] across={across} a1={a1} a2={a2} r={r:.6f}", flush=True)
    fix.post_force(0, f)
    assert abs(fix.cv("D.dist") - r) <= 1e-14 * r
    assert abs(fix.cv("Q6.AVE") - q6) <= 1e-5 * q6
    wd, w6 = fix.cv_weight("D.dist"), fix.cv_weight("Q6.AVE")
    assert wd != 0.0 and w6 != 0.0
    F_ref = -env.dQ_dx(6, w6)
    F_ref[a1] += wd * d / r          # -dV/dr * dr/dx_a1 with dr/dx_a1 = -d/r
    F_ref[a2] -= wd * d / r
    assert rel_force_err(f, F_ref) <= F_TOL
    # the distance part is exact: the pair's force difference from the Q6 part
    Fq = -env.dQ_dx(6, w6)
    assert np.abs((f[a1] - Fq[a1]) - wd * d / r).max() <= 1e-4 * abs(wd)
    fix.close()
    ctx.close()


def test_distance_in_walkers_and_graph_runs(mdt, orc):
    """Three walkers, one graph launch per batch of steps: each walker's distance is its own."""
    systems = [fcc_system(mdt, 4, sigma=0.1, seed=82 + k) for k in range(3)]
    ctx = mdt.Context(0, 3)
    for w, sm in enumerate(systems):
        ctx.bind_system(w, sm)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.05 0.005 10 PACE 2 HILLS_FILE NONE CAL NAME D DISTANCE 3 40 "
                       "CV_dim 1 SYMBOL v1 D.AVE DIM 1 0 20 400 v1 WT 1")
    fix.zero_forces()
    fix.run(0, 9)
    fix.sync()
    for w, sm in enumerate(systems):
        x = sm.arrays()[0][: sm.n_local]
        lo, hi = sm.box()
        r, _ = orc.distance_cv(x, np.asarray(hi) - np.asarray(lo), 2, 39)
        assert abs(fix.cv("D.AVE", w) - r) <= 1e-14 * r
        F = fix.forces(sm.n_local, w)
        moved = np.flatnonzero(np.abs(F).sum(axis=1))
        assert set(moved.tolist()) <= {2, 39}
        assert np.abs(F[2] + F[39]).max() <= 1e-15 * max(1.0, np.abs(F).max())
    fix.close()
    ctx.close()


def test_distance_atom_beyond_a_walker_is_refused(mdt):
    systems = [fcc_system(mdt, 4, sigma=0.1, seed=85), fcc_system(mdt, 3, sigma=0.1, seed=86)]   # 256, 108 atoms
    ctx = mdt.Context(0, 2)
    for w, sm in enumerate(systems):
        ctx.bind_system(w, sm)
    fix = mdt.Fix(ctx, "GAUSSIAN 0.05 0.005 10 PACE 1 HILLS_FILE NONE CAL NAME D DISTANCE 3 200 "
                       "CV_dim 1 SYMBOL v1 D.AVE DIM 1 0 20 400 v1 WT 0")
    with pytest.raises(mdt.MdtError) as e:
        fix.init()
    assert "beyond walker 1" in str(e.value)
    ctx.close()
