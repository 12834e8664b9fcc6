"""STRU_FACTOR (Debye form, src/cvs/CVs_Strufactor.cu) on the GPU against the oracle
(run with -m gpu).

Bars: the CV within 1e-12 of the oracle relative (FP64 throughout; sin/cos may differ in the
last bit); forces within 1e-9 relative to their maximum (max |F - F_ref| / max |F_ref|).
"""
import numpy as np
import pytest

from helpers import fcc_system, rel_force_err

pytestmark = pytest.mark.gpu

RC, Q = 5.5, 2.0


@pytest.mark.parametrize("grouped", [False, True])
def test_stru_factor_value_and_forces(mdt, orc, grouped):
    s = fcc_system(mdt, 5, sigma=0.15, seed=91)
    x, off, idx, own = s.arrays()
    n = s.n_local
    grp = (x[:n, 2] < x[:n, 2].mean()).astype(np.uint8) if grouped else None
    cv = orc.sf_cv(x, off, idx, n, RC, Q, grp)
    gname = "lower" if grouped else "all"
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    if grouped:
        ctx.bind_group(0, grp)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.05 0.005 10 PACE 0 HILLS_FILE NONE "
                       f"CAL NAME S STRU_FACTOR {gname} cutoff_r {RC} q_factor {Q} "
                       "CV_dim 1 SYMBOL v1 S.AVE DIM 1 0 5 500 v1 WT 0")
    fix.init()
    fix.add_hills(np.array([[cv + 0.02]]), np.array([0.005]), 0.05)
    f = np.zeros((n, 3))
    print(f"[strufactor] grouped={grouped} cv={cv:.8f}", flush=True)
    fix.post_force(0, f)
    assert abs(fix.cv("S.AVE") - cv) <= 1e-12 * abs(cv)
    w = fix.cv_weight("S.AVE")
    assert w != 0.0
    F_ref = -w * orc.sf_grad(x, off, idx, n, own, RC, Q, grp)
    assert rel_force_err(f, F_ref) <= 1e-9
    if grouped:
        # atoms outside the group that neighbour it carry force too (the exact gradient)
        outside = ~grp.astype(bool)
        assert np.abs(f[outside]).max() > 0.0
    fix.close()
    ctx.close()


@pytest.mark.parametrize("swname", ["FERMI", "RATIONAL", "TANH"])
def test_stru_factor_count_with_switching_function(mdt, orc, swname):
    """S.COUNT = sum over the group of sw(S_i) (compute_cv_COUNT), with its exact forces."""
    spec, sw = {"FERMI": ("FERMI r_0 2.0 alpha 4", dict(type="FERMI", r0=2.0, alpha=4.0)),
                "RATIONAL": ("RATIONAL r_0 2.0 d_0 0.5 n 6 m 12", dict(type="RATIONAL", r0=2.0, d0=0.5, n=6, m=12)),
                "TANH": ("TANH r_0 2.1 alpha 3", dict(type="TANH", r0=2.1, alpha=3.0))}[swname]
    s = fcc_system(mdt, 5, sigma=0.15, seed=92)
    x, off, idx, own = s.arrays()
    n = s.n_local
    grp = (x[:n, 0] < x[:n, 0].mean()).astype(np.uint8)
    cv = orc.sf_count(x, off, idx, n, RC, Q, sw, grp)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    ctx.bind_group(0, grp)
    fix = mdt.Fix(ctx, f"GAUSSIAN 2.0 0.005 10 PACE 0 HILLS_FILE NONE "
                       f"CAL NAME S STRU_FACTOR left cutoff_r {RC} q_factor {Q} SW_func {spec} "
                       "CV_dim 1 SYMBOL v1 S.COUNT DIM 1 0 600 600 v1 WT 0")
    fix.init()
    fix.add_hills(np.array([[cv + 1.0]]), np.array([0.005]), 2.0)
    f = np.zeros((n, 3))
    print(f"[strufactor count] {swname} cv={cv:.8f}", flush=True)
    fix.post_force(0, f)
    assert abs(fix.cv("S.COUNT") - cv) <= 1e-12 * abs(cv)
    w = fix.cv_weight("S.COUNT")
    assert w != 0.0
    F_ref = -w * orc.sf_count_grad(x, off, idx, n, own, RC, Q, sw, grp)
    assert rel_force_err(f, F_ref) <= 1e-9
    fix.close()
    ctx.close()


@pytest.mark.parametrize("kind,func", [("STRU_FACTOR", "AVE"), ("WEIGHT_CHEM", "AVE"), ("WEIGHT_CHEM", "COUNT")])
def test_chemically_weighted_pairs(mdt, orc, kind, func):
    """STRU_FACTOR ... Chemical (CVs_Strufactor_chem.cu) and WEIGHT_CHEM
    (CV_Weighted_Chemical_Pair_Order.cu): every pair term carries the two atoms' weights,
    a Chem_map value for mapped types and exp(-(t - c)^2 / (2 sigma^2)) for the rest; the
    types come from mdt_bind_types, ghosts take their owner's."""
    s = fcc_system(mdt, 5, sigma=0.15, seed=93)
    x, off, idx, own = s.arrays()
    n = s.n_local
    types = (np.random.default_rng(94).integers(1, 4, n)).astype(np.int32)
    t_all = np.concatenate([types, types[own]])
    aw = orc.chem_weights(t_all, c_target=2.0, sigma=0.7, cmap={3: 0.45})
    grp = (x[:n, 1] < x[:n, 1].mean()).astype(np.uint8)
    pair = 1 if kind == "WEIGHT_CHEM" else 0
    chem = "Chem_ctarget 2.0 Chem_sigma 0.7 Chem_map (3 0.45)"
    if kind == "STRU_FACTOR":
        cal = f"CAL NAME S STRU_FACTOR left cutoff_r {RC} q_factor {Q} Chemical {chem}"
    else:
        cal = f"CAL NAME S WEIGHT_CHEM left cutoff_r {RC} {chem}"
    sw = None
    if func == "COUNT":
        sw = dict(type="RATIONAL", r0=3.0, d0=0.5, n=6, m=12)
        cal += " SW_func RATIONAL r_0 3.0 d_0 0.5 n 6 m 12"
        cv = orc.sf_count(x, off, idx, n, RC, Q, sw, grp, aw, pair)
        dim = "DIM 1 0 600 600"
    else:
        cv = orc.sf_cv(x, off, idx, n, RC, Q, grp, aw, pair)
        dim = "DIM 1 0 20 2000"
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    ctx.bind_group(0, grp)
    ctx.bind_types(0, types)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.5 0.005 10 PACE 0 HILLS_FILE NONE {cal} "
                       f"CV_dim 1 SYMBOL v1 S.{func} {dim} v1 WT 0")
    fix.init()
    fix.add_hills(np.array([[cv + 0.1]]), np.array([0.005]), 0.5)
    f = np.zeros((n, 3))
    print(f"[chem] {kind} {func} cv={cv:.8f}", flush=True)
    fix.post_force(0, f)
    assert abs(fix.cv(f"S.{func}") - cv) <= 1e-12 * abs(cv)
    w = fix.cv_weight(f"S.{func}")
    assert w != 0.0
    if func == "COUNT":
        G = orc.sf_count_grad(x, off, idx, n, own, RC, Q, sw, grp, aw, pair)
    else:
        G = orc.sf_grad(x, off, idx, n, own, RC, Q, grp, aw, pair)
    assert rel_force_err(f, -w * G) <= 1e-9
    fix.close()
    ctx.close()


def test_chemical_cal_without_types_is_refused(mdt):
    s = fcc_system(mdt, 3, sigma=0.05, seed=95)
    ctx = mdt.Context(0, 1)
    ctx.bind_system(0, s)
    fix = mdt.Fix(ctx, f"GAUSSIAN 0.5 0.005 10 PACE 0 HILLS_FILE NONE CAL NAME S WEIGHT_CHEM all cutoff_r {RC} "
                       "CV_dim 1 SYMBOL v1 S.AVE DIM 1 0 20 200 v1 WT 0")
    with pytest.raises(mdt.MdtError, match="bind the atom types"):
        fix.init()
    fix.close()
    ctx.close()
