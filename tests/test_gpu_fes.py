"""FES reconstruction on the GPU (mdt_fes_sum, mdt_fes_marginal, python -m mdt.fes) against the
oracle's numpy restatement of tools/plot_hills.py (run with -m gpu).

Bars: the hill sums within 1e-13 of the oracle relative to their maximum (the same FP64
operations in the same hill order; only exp may differ in the last bit); marginals within
1e-10 kT-scaled (the oracle's numpy sum over a row is pairwise, the kernel's sequential).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KBT = 8.3144621 * 300.0 / 1000.0
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _hills_2d(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([np.arange(n) * 10.0, rng.uniform(0.1, 0.9, n), rng.uniform(-0.4, 0.4, n),
                            rng.uniform(0.001, 0.005, n), np.full(n, 0.03)])


def test_fes_1d_reference_hills(mdt, orc):
    """The reference's own HILLS fixture (238 1-D hills) on a 1000-bin grid over [0, 1]."""
    rows = np.loadtxt(os.path.join(GOLDEN, "HILLS_reference_test_cvset"), comments="#")
    x = np.linspace(0.0005, 0.9995, 1000)
    ctx = mdt.Context(0, 1)
    F = mdt.fes_sum(ctx, rows, x)
    ref = orc.fes_sum_1d(rows, x)
    assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max()
    ctx.close()


def test_fes_2d_windows_and_marginals(mdt, orc):
    """2-D windows (hills at the grid edges included) and both marginals."""
    rows = _hills_2d(400, 31)
    rows[:5, 1] = [0.0, 1.0, 0.002, 0.999, 0.5]          # windows clipped at either edge
    rows[:5, 2] = [-0.5, 0.5, 0.0, -0.499, 0.4999]
    lo, hi, nb = np.array([0.0, -0.5]), np.array([1.0, 0.5]), np.array([120, 90])
    d = (hi - lo) / nb
    x = np.linspace(lo[0] + d[0] / 2, hi[0] - d[0] / 2, nb[0])
    y = np.linspace(lo[1] + d[1] / 2, hi[1] - d[1] / 2, nb[1])
    ctx = mdt.Context(0, 1)
    F = mdt.fes_sum(ctx, rows, x, y, lo, d)
    ref = orc.fes_sum_2d(rows, x, y, lo, d)
    assert F.shape == (120, 90)
    assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max()
    for axis, w in ((1, d[1]), (0, d[0])):
        m = mdt.fes_marginal(ctx, F, axis, KBT, w)
        mref = orc.fes_marginal(ref, KBT, axis, w)
        assert m.min() == 0.0
        assert np.abs(m - mref).max() <= 1e-10 * KBT
    ctx.close()


def test_fes_command_line(mdt, orc, tmp_path, monkeypatch):
    """python -m mdt.fes <input> <HILLS> writes FES.dat and, for two CVs, the marginals."""
    from mdt import fes
    rows = _hills_2d(150, 32)
    hills = tmp_path / "HILLS2"
    with open(hills, "w") as f:
        f.write("# step cv1 cv2 height sigma\n")
        for r in rows:
            f.write(f"{int(r[0])} {r[1]:.16g} {r[2]:.16g} {r[3]:.16g} {r[4]:.16g}\n")
    deck = tmp_path / "run.in"
    deck.write_text("fix 1 all metad GAUSSIAN 0.03 0.005 10 PACE 10 CAL NAME Q6 STEINH Q 6 all "
                    "CAL NAME Q4 STEINH Q 4 all CV_dim 2 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE "
                    "DIM 1 0 1 100 v1 DIM 2 -0.5 0.5 80 v2 WT 1\n")
    monkeypatch.chdir(tmp_path)
    assert fes.main([str(deck), str(hills)]) == 0
    F = np.loadtxt(tmp_path / "FES.dat")
    d = np.array([0.01, 1.0 / 80])
    x = np.linspace(0.005, 0.995, 100)
    y = np.linspace(-0.5 + d[1] / 2, 0.5 - d[1] / 2, 80)
    ref = orc.fes_sum_2d(np.loadtxt(hills, comments="#"), x, y, [0.0, -0.5], d)
    assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max()
    m1 = np.loadtxt(tmp_path / "Marginal_FES_CV1.dat", comments="#")
    assert np.abs(m1[:, 0] - x).max() <= 1e-8 * np.abs(x).max()
    assert np.abs(m1[:, 1] - orc.fes_marginal(ref, KBT, 1, d[1])).max() <= 1e-7 * KBT
    assert os.path.exists(tmp_path / "Marginal_FES_CV2.dat")


def test_fes_without_hills(mdt):
    ctx = mdt.Context(0, 1)
    F = mdt.fes_sum(ctx, np.zeros((0, 4)), np.linspace(0.0, 1.0, 11))
    assert F.shape == (11,) and not F.any()
    ctx.close()
