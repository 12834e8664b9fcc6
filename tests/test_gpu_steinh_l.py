"""STEINH L (switching-weighted local Q, src/cvs/CVs_Steinhardt_Local.cu) on the GPU against
the oracle (run with -m gpu).

Bars: the CV within 1e-12 of the oracle relative (FP64 throughout; the per-tile sums and the
Y_lm polynomial differ from the oracle's in rounding only); forces within 1e-9 relative to
their maximum (max |F - F_ref| / max |F_ref|), the oracle's gradient being exact and pinned by
finite differences in test_oracle.py.
"""
import time

import numpy as np
import pytest

from helpers import fcc_system, rel_force_err

pytestmark = pytest.mark.gpu

SWR = ("RATIONAL r_0 3.0", dict(type="RATIONAL", r0=3.0, d0=0.0, n=6, m=12))
SWC = ("FERMI r_0 0.35 alpha 10", dict(type="FERMI", r0=0.35, alpha=10.0))
GRID = {"MEAN_SOLID": ("0.05", "0 5 500", 0.02), "FRAC": ("0.05", "0 5 500", 0.02),
        "NSOLID": ("2.0", "0 600 600", 1.0)}


def _deck(l, mode, grp, swr=SWR[0], swc=SWC[0], eps=1e-6):
    sig, dim, _ = GRID[mode]
    return (f"GAUSSIAN {sig} 0.005 10 PACE 0 HILLS_FILE NONE "
            f"CAL NAME sr SW_FUNC {swr} CAL NAME sc SW_FUNC {swc} "
            f"CAL NAME LQ STEINH L {l} {grp} cutoff_eps {eps} SW_FUNC_r sr SW_FUNC_cv sc "
            f"CV_dim 1 SYMBOL v1 LQ.{mode} DIM 1 {dim} v1 WT 0")


@pytest.mark.parametrize("mode", ["MEAN_SOLID", "NSOLID", "FRAC"])
@pytest.mark.parametrize("l,grouped", [(6, False), (6, True), (4, True), (3, False)])
def test_steinh_l_value_and_forces(mdt, orc, mode, l, grouped):
    s = fcc_system(mdt, 4, sigma=0.35, sigma_above=0.1, seed=101 + l)
    x, off, idx, own = s.arrays()
    n = s.n_local
    grp = (x[:n, 2] < x[:n, 2].mean()).astype(np.uint8) if grouped else None
    cv = orc.lq_cv(x, off, idx, n, own, l, SWR[1], SWC[1], 1e-6, mode, grp)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    if grouped:
        ctx.bind_group(0, grp)
    fix = mdt.Fix(ctx, _deck(l, mode, "lower" if grouped else "all"))
    fix.init()
    sig = float(GRID[mode][0])
    fix.add_hills(np.array([[cv + GRID[mode][2]]]), np.array([0.005]), sig)
    f = np.zeros((n, 3))
    print(f"[steinh L] l={l} mode={mode} grouped={grouped} cv={cv:.10f}", flush=True)
    fix.post_force(0, f)
    got = fix.cv(f"LQ.{mode}")
    assert abs(got - cv) <= 1e-12 * abs(cv), (got, cv)
    w = fix.cv_weight(f"LQ.{mode}")
    assert w != 0.0
    F_ref = -w * orc.lq_grad(x, off, idx, n, own, l, SWR[1], SWC[1], 1e-6, mode, grp)
    assert rel_force_err(f, F_ref) <= 1e-9
    assert np.abs(f.sum(axis=0)).max() <= 1e-10 * np.abs(f).max() * n     # periodic: sum F = 0
    if grouped:
        assert np.abs(f[~grp.astype(bool)]).max() > 0.0      # atoms outside the group carry force
    fix.close()
    ctx.close()


def test_steinh_l_step_weights_and_eps(mdt, orc):
    """STEP weights (a hard shell) and a cutoff_eps that drops the tail of a TANH sigma."""
    s = fcc_system(mdt, 4, sigma=0.3, seed=111)
    x, off, idx, own = s.arrays()
    n = s.n_local
    line = dict(type="LINE")
    for spec, sw, eps in [("STEP r_0 4.2", dict(type="STEP", r0=4.2), 1e-6),
                          ("TANH r_0 3.6 alpha -3", dict(type="TANH", r0=3.6, alpha=-3.0), 0.02)]:
        cv = orc.lq_cv(x, off, idx, n, own, 6, sw, line, eps, "MEAN_SOLID")
        ctx = mdt.Context(0, 1)
        ctx.bind_system(0, s)
        deck = (f"GAUSSIAN 0.05 0.005 10 PACE 0 HILLS_FILE NONE CAL NAME sr SW_FUNC {spec} "
                f"CAL NAME LQ STEINH L 6 all cutoff_eps {eps} SW_FUNC_r sr "
                "CV_dim 1 SYMBOL v1 LQ.AVE DIM 1 0 5 500 v1 WT 0")
        fix = mdt.Fix(ctx, deck)
        fix.init()
        fix.add_hills(np.array([[cv + 0.02]]), np.array([0.005]), 0.05)
        f = np.zeros((n, 3))
        print(f"[steinh L] sigma={spec} eps={eps} cv={cv:.10f}", flush=True)
        fix.post_force(0, f)
        assert abs(fix.cv("LQ.AVE") - cv) <= 1e-12 * cv
        F_ref = -fix.cv_weight("LQ.AVE") * orc.lq_grad(x, off, idx, n, own, 6, sw, line, eps, "MEAN_SOLID")
        assert rel_force_err(f, F_ref) <= 1e-9
        fix.close()
        ctx.close()


def test_steinh_l_walkers_in_graph_runs(mdt, orc):
    """Two walkers of different sizes, one graph launch per batch of steps: each walker's CV
    is its own system's."""
    systems = [fcc_system(mdt, 4, sigma=0.2, seed=120), fcc_system(mdt, 3, sigma=0.4, seed=121)]
    ctx = mdt.Context(0, 2)
    for w, sm in enumerate(systems):
        ctx.bind_system(w, sm)
    fix = mdt.Fix(ctx, _deck(6, "FRAC", "all").replace("PACE 0", "PACE 3").replace("WT 0", "WT 1"))
    fix.zero_forces()
    fix.run(0, 7)
    fix.sync()
    for w, sm in enumerate(systems):
        x, off, idx, own = sm.arrays()
        ref = orc.lq_cv(x, off, idx, sm.n_local, own, 6, SWR[1], SWC[1], 1e-6, "FRAC")
        print(f"[steinh L walkers] w={w} cv={fix.cv('LQ.FRAC', w):.12f} ref={ref:.12f}", flush=True)
        assert abs(fix.cv("LQ.FRAC", w) - ref) <= 1e-12 * ref
        F = fix.forces(sm.n_local, w)
        assert np.abs(F).max() > 0.0
        assert np.abs(F.sum(axis=0)).max() <= 1e-10 * np.abs(F).max() * sm.n_local
    fix.close()
    ctx.close()


def test_steinh_l_256k_perfect_fcc_known_answer(mdt):
    """256,000-atom perfect FCC, STEP r_0 = 3.9: every atom's LQ_6 is the 12-neighbour q_6 =
    0.574524, so MEAN_SOLID = 0.574524 and, biased, the forces vanish with the gradient's
    symmetry; then a jittered system of the same size: the forces sum to zero.  Reports the
    time per step."""
    s = fcc_system(mdt, 40, sigma=0.0)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    deck = ("GAUSSIAN 0.05 0.005 10 PACE 0 HILLS_FILE NONE CAL NAME sr SW_FUNC STEP r_0 3.9 "
            "CAL NAME LQ STEINH L 6 all SW_FUNC_r sr CV_dim 1 SYMBOL v1 LQ.AVE DIM 1 0 2 2000 v1 WT 0")
    fix = mdt.Fix(ctx, deck)
    fix.init()
    fix.add_hills(np.array([[0.6]]), np.array([0.005]), 0.05)
    f = np.zeros((s.n_local, 3))
    fix.post_force(0, f)
    print(f"[steinh L 256k] perfect FCC cv={fix.cv('LQ.AVE'):.9f} max|F|={np.abs(f).max():.3e}", flush=True)
    assert abs(fix.cv("LQ.AVE") - 0.574524) <= 1e-6
    assert np.abs(f).max() <= 1e-12
    fix.close()
    ctx.close()
    s = fcc_system(mdt, 40, sigma=0.6, sigma_above=0.1, seed=2024)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, _deck(6, "MEAN_SOLID", "all").replace("PACE 0", "PACE 10"))
    fix.zero_forces()
    fix.run(0, 10)
    fix.sync()
    t0 = time.perf_counter()
    fix.run(10, 50)
    fix.sync()
    dt = (time.perf_counter() - t0) / 50
    F = fix.forces(s.n_local, 0)
    print(f"[steinh L 256k] melted half: cv={fix.cv('LQ.MEAN_SOLID'):.9f} {1e6 * dt:.1f} us/step "
          f"{s.n_local / dt:.3e} atom-steps/s", flush=True)
    assert np.abs(F).max() > 0.0
    assert np.abs(F.sum(axis=0)).max() <= 1e-9 * np.abs(F).max() * s.n_local
    fix.close()
    ctx.close()


def test_steinh_l_on_device_built_lists(mdt):
    """Lists from the device neighbour build give the same bits as the host's (same order,
    so the same per-tile ELL layout and sums)."""
    s = fcc_system(mdt, 5, sigma=0.3, sigma_above=0.1, seed=131)
    n = s.n_local
    out = []
    for use_device in (False, True):
        ctx = mdt.Context(0, 1)
        if use_device:
            lo, hi = s.box()
            ctx.build_neighbors(s.arrays()[0][:n], lo, hi, 5.9)
            ctx.bind_built_neighbors(0)
        else:
            ctx.bind_system(0, s)
        fix = mdt.Fix(ctx, _deck(6, "NSOLID", "all").replace("PACE 0", "PACE 1"))
        fix.init()
        f = np.zeros((n, 3))
        for step in range(3):
            fix.post_force(step, f)
        out.append((f.copy(), fix.cv("LQ.NSOLID")))
        fix.close()
        ctx.close()
    print(f"[steinh L device lists] cv={out[1][1]:.12f} max|F|={np.abs(out[1][0]).max():.3e}", flush=True)
    assert np.abs(out[0][0]).max() > 0.0
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_steinh_l_256k_half_melted_against_c_oracle(mdt, orc):
    """The 256,000-atom half-melted FCC system (40^3 cells, lower half jittered 0.6 A) under
    STEINH L 6 MEAN_SOLID, value and forces against the OpenMP C restatement of the oracle
    (itself equal to the Python STEINH L oracle to 1e-12, test_oracle.py)."""
    from mdt.workloads import c2_system
    s = c2_system(mdt)
    x, off, idx, own = s.arrays()
    n = s.n_local
    t0 = time.perf_counter()
    cv, G = orc.lq_c(x, off, idx, n, own, 6, SWR[1], SWC[1], 1e-6, "MEAN_SOLID")
    print(f"[steinh L 256k] oracle {time.perf_counter() - t0:.1f} s, cv={cv:.10f}", flush=True)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, _deck(6, "MEAN_SOLID", "all"))
    fix.init()
    fix.add_hills(np.array([[cv + 0.02]]), np.array([0.005]), 0.05)
    f = np.zeros((n, 3))
    fix.post_force(0, f)
    got = fix.cv("LQ.MEAN_SOLID")
    assert abs(got - cv) <= 1e-12 * abs(cv), (got, cv)
    w = fix.cv_weight("LQ.MEAN_SOLID")
    assert w != 0.0
    assert rel_force_err(f, -w * G) <= 1e-9
    fix.close()
    ctx.close()


@pytest.mark.parametrize("mode", ["MEAN_SOLID", "FRAC"])
def test_steinh_l_dilute_gas_short_and_empty_lists(mdt, orc, mode):
    """A dilute random gas (about 3 list entries per atom within 5.9 A): many atoms have empty,
    one- or two-entry lists, so the gathers' software pipelines start, run and end on lists
    shorter than their depth, and isolated atoms take no part.  CV and forces follow the
    oracle to the FP64 bars."""
    L = 83.0
    s = mdt.System().set_box([0, 0, 0], [L, L, L]).add_uniform(2000, 0.0, L, seed=141).build(5.9)
    x, off, idx, own = s.arrays()
    n = s.n_local
    lens = np.diff(off[: n + 1])
    assert (lens == 0).sum() > 50 and (lens == 1).sum() > 100 and (lens == 2).sum() > 100, np.bincount(lens)
    cv = orc.lq_cv(x, off, idx, n, own, 6, SWR[1], SWC[1], 1e-6, mode, None)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, _deck(6, mode, "all"))
    fix.init()
    fix.add_hills(np.array([[cv + GRID[mode][2]]]), np.array([0.005]), float(GRID[mode][0]))
    f = np.zeros((n, 3))
    print(f"[steinh L dilute] mode={mode} n={n} mean list {lens.mean():.2f} cv={cv:.10f}", flush=True)
    fix.post_force(0, f)
    got = fix.cv(f"LQ.{mode}")
    assert abs(got - cv) <= 1e-12 * abs(cv), (got, cv)
    w = fix.cv_weight(f"LQ.{mode}")
    assert w != 0.0
    F_ref = -w * orc.lq_grad(x, off, idx, n, own, 6, SWR[1], SWC[1], 1e-6, mode, None)
    assert np.abs(F_ref).max() > 0.0
    assert rel_force_err(f, F_ref) <= 1e-9
    assert np.abs(f[lens == 0]).max() == 0.0          # isolated atoms feel no bias force
    fix.close()
    ctx.close()
