"""METAD_RESTART on the GPU against the reference's HILLS format (GH_t0_uniformGrid.cu:100-192).

- The reference's own HILLS fixture (test/test_cvset/HILLS, 238 1-D hills) replayed into the
  C1 grid equals the committed golden grid.
- A 2-D HILLS file whose hills carry sigmas different from GAUSSIAN's is replayed hill by
  hill with each hill's own sigma, as the reference's add_to_grid(cv, h, s) does.
- A hills-list bias (Gaussian_Hill_type 2) keeps every replayed hill, well past the
  initial 4,096-hill log, each with its own sigma.
Bars: grids 1e-12 of their maximum (the same FP64 arithmetic); hills sums 1e-12 relative.
"""
import os
import shutil

import numpy as np
import pytest

from helpers import fcc_system

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
Q6 = "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
Q4 = "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "


def _ctx(mdt, s):
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    return ctx


def test_restart_from_reference_hills_fixture(mdt, tmp_path):
    s = fcc_system(mdt, 6, sigma=0.1, seed=3)
    path = str(tmp_path / "HILLS")
    shutil.copy(os.path.join(GOLD, "HILLS_reference_test_cvset"), path)
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE {path} " + Q6 +
                       "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0 METAD_RESTART 1")
    fix.init()
    gold = np.load(os.path.join(GOLD, "grid_reference_hills.npy"))
    g = fix.grid(1000)
    print(f"[restart] reference fixture: grid max {g.max():.6f}", flush=True)
    assert np.abs(g - gold).max() <= 1e-12 * np.abs(gold).max()
    # the run appends to the file; the replayed hills are not written again
    fix.post_force(0)
    fix.close()
    lines = [l for l in open(path) if not l.startswith("#")]
    assert len(lines) == 239 and lines[-1].split()[0] == "0"
    ctx.close()


def test_restart_2d_hills_with_their_own_sigma(mdt, orc, tmp_path):
    rng = np.random.default_rng(12)
    nh = 300
    c = np.column_stack([rng.uniform(0.1, 0.9, nh), rng.uniform(0.05, 0.6, nh)])
    h = rng.uniform(0.001, 0.005, nh)
    sg = rng.choice([0.01, 0.015, 0.025], nh)          # GAUSSIAN's sigma is 0.01
    path = str(tmp_path / "HILLS")
    with open(path, "w") as f:
        f.write("# step cv1 cv2 height sigma\n")
        for k in range(nh):
            f.write(f"{10 * k} {c[k, 0]:.17g} {c[k, 1]:.17g} {h[k]:.17g} {sg[k]:.17g}\n")
    s = fcc_system(mdt, 6, sigma=0.1, seed=3)
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE {path} " + Q6 + Q4 +
                       "CV_dim 2 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE DIM 1 0 1 200 v1 DIM 2 0 1 150 v2 "
                       "WT 0 METAD_RESTART 1")
    fix.init()
    g = orc.Grid([0.0, 0.0], [1.0, 1.0], [200, 150])
    for k in range(nh):
        g.deposit(c[k], h[k], sg[k])
    got = fix.grid(200 * 150)
    assert np.abs(got - g.g).max() <= 1e-12 * np.abs(g.g).max()
    # the same file replayed with every sigma taken as GAUSSIAN's would differ
    g1 = orc.Grid([0.0, 0.0], [1.0, 1.0], [200, 150])
    for k in range(nh):
        g1.deposit(c[k], h[k], 0.01)
    assert np.abs(got - g1.g).max() > 1e-3 * np.abs(g.g).max()
    fix.close()
    ctx.close()


def test_restart_hills_list_beyond_initial_log(mdt, orc, tmp_path):
    """Gaussian_Hill_type 2 (exact hills sum) restarted from 5,000 hills with two sigmas."""
    rng = np.random.default_rng(13)
    nh = 5000
    c = rng.uniform(0.2, 0.8, nh)
    h = rng.uniform(0.001, 0.005, nh)
    sg = np.where(rng.random(nh) < 0.3, 0.02, 0.01)
    path = str(tmp_path / "HILLS")
    with open(path, "w") as f:
        f.write("# step cv_values height sigma\n")
        for k in range(nh):
            f.write(f"{k} {c[k]:.17g} {h[k]:.17g} {sg[k]:.17g}\n")
    s = fcc_system(mdt, 6, sigma=0.1, seed=3)
    ctx = _ctx(mdt, s)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE {path} " + Q6 +
                       "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0 Gaussian_Hill_type 2 METAD_RESTART 1")
    fix.init()
    rows = fix.hills(1)
    assert rows.shape[0] == nh
    assert np.array_equal(rows[:, 1], c) and np.array_equal(rows[:, 3], sg)
    pts = np.linspace(0.1, 0.9, 37)[:, None]
    V, dV = fix.bias_eval(pts)
    Vr, dVr = orc.hills_sum(c[:, None], h, [0.01], pts, hill_sigma=sg)
    assert np.abs(V - Vr).max() <= 1e-12 * np.abs(Vr).max()
    assert np.abs(dV - dVr).max() <= 1e-12 * np.abs(dVr).max()
    fix.close()
    ctx.close()
