"""CPU tests of the host side: system builder, neighbour lists, C-ABI exports, fix parser."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import fcc_system, liquid_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(mdt):
    hdr = open(os.path.join(ROOT, "include", "mdt.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(mdt_\w+)\s*\(", hdr, re.M))
    assert len(names) > 40
    L = ctypes.CDLL(mdt.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert mdt.lib().mdt_version().decode().startswith("mdt")


def test_fcc_generator_counts_and_box(mdt):
    s = mdt.System().add_fcc(3, 4, 5, 5.26)
    lo, hi = s.box()
    assert np.allclose(hi - lo, [3 * 5.26, 4 * 5.26, 5 * 5.26])
    assert s.n_local == 4 * 60
    x = s.arrays()[0]
    assert x.min() >= 0 and np.all(x < hi)


def test_jitter_is_seeded_and_split(mdt):
    a = mdt.System().add_fcc(6, 6, 6, 5.26, 0.6, 0.1, 3 * 5.26, seed=4).arrays()[0]
    b = mdt.System().add_fcc(6, 6, 6, 5.26, 0.6, 0.1, 3 * 5.26, seed=4).arrays()[0]
    p = mdt.System().add_fcc(6, 6, 6, 5.26).arrays()[0]
    assert np.array_equal(a, b)
    L = 6 * 5.26
    d = (a - p + L / 2) % L - L / 2
    low = p[:, 2] < 3 * 5.26
    assert 0.5 < d[low].std() < 0.7 and 0.08 < d[~low].std() < 0.12


@pytest.mark.parametrize("kind", ["fcc", "liquid"])
def test_neighbour_list_bruteforce(mdt, kind):
    s = fcc_system(mdt, 4, sigma=0.3, seed=2) if kind == "fcc" else liquid_system(mdt, 800, seed=1)
    x, off, idx, own = s.arrays()
    n = s.n_local
    lo, hi = s.box()
    L = hi - lo
    rg = 5.9
    # ghosts are exact images of their owners and lie within rg of the box
    shift = x[n:] - x[own]
    assert np.allclose(shift / L - np.round(shift / L), 0, atol=1e-9)
    assert np.all(x[n:] >= lo - rg) and np.all(x[n:] < hi + rg)
    # every local pair within rg under minimum image appears exactly once per center
    xl = x[:n]
    for i in range(0, n, 7):
        d = xl - xl[i]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=1)
        expect = sorted(j for j in np.nonzero(r2 <= rg * rg)[0] if j != i)
        got = idx[off[i]:off[i + 1]]
        owners = np.where(got < n, got, own[np.clip(got - n, 0, None)])
        assert sorted(owners.tolist()) == expect
        dd = x[got] - x[i]
        assert np.all(np.sum(dd * dd, axis=1) <= rg * rg)


def test_box_too_small_is_rejected(mdt):
    s = mdt.System().add_fcc(2, 2, 2, 5.26)
    with pytest.raises(mdt.MdtError):
        s.build(5.9)


GOOD = ("GAUSSIAN 0.01 0.005 10 PACE 10 "
        "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
        "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
        "CV_dim 1 SIMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE DIM 1 0 1 1000 \"v1+v2\" METAD_RESTART 0 WT 0")


def test_fix_parser_accepts_reference_syntax(mdt):
    ctx = mdt.Context(-1, 1)
    fix = mdt.Fix(ctx, GOOD)
    fix.close()
    # README example, RECORD block included
    mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10.0 PACE 10 RECORD FILE_NAME rec.log REC_PACE 1 "
                 "CAL NAME Q6 STEINH Q 6 metad_group cutoff_r 3.9 cutoff_Natoms 18 "
                 "CAL NAME Q4 STEINH Q 4 metad_group cutoff_r 3.9 cutoff_Natoms 18 "
                 "CV_dim 1 SIMBOL v1 Q6.AVE SIMBOL v2 Q4.AVE DIM 1 0 1 1000 v1+v2 METAD_RESTART 1 WT 0").close()
    # expressions with functions and constants compile
    mdt.Fix(ctx, GOOD.replace('"v1+v2"', '"sqrt(v1^2+v2^2)*cos(pi/4)+max(v1,v2)-exp(-v1)"')).close()


@pytest.mark.parametrize("bad,code", [
    ("GAUSSIAN 0.01 0.005", -22),
    ("PACE", -22),
    ("FOO 1", -22),
    ("CAL NAME Q6 STEINH Q 5 all", -22),
    ("CAL NAME Q6 STEINH Q 8 all", -95),
    ("CAL NAME Q6 STEINH L 8 all", -95),
    ("CAL NAME Q6 STEINH Q 6 all SW_FUNC_r sw", -22),
    ("CAL NAME sw SW_FUNC NOPE", -22),
    ("CAL NAME sw SW_FUNC RATIONAL r_0 0", -22),
    ("CAL NAME sw SW_FUNC STEP CAL NAME sw SW_FUNC FERMI", -22),
    ("CAL NAME Q6 WEIGHT_CHEM 1 2", -22),
    ("CAL NAME W WEIGHT_CHEM all Chem_map (1", -22),
    ("CAL NAME W WEIGHT_CHEM all Chem_map (1 2", -22),
    ("CAL NAME W WEIGHT_CHEM all Chem_map (1)", -22),
    ("CAL NAME S STRU_FACTOR all cutoff_r 5", -22),
    ("CAL NAME S STRU_FACTOR all q_factor 2 SW_func NOPE", -22),
    ("CAL NAME S STRU_FACTOR all q_factor 0", -22),
    ("CAL NAME D DISTANCE 3 3", -22),
    ("CAL NAME D DISTANCE 0 2", -22),
    ("CAL NAME D DISTANCE 4", -22),
    ("CAL NAME Q6 NOPE 1 2", -22),
])
def test_fix_parser_errors(mdt, bad, code):
    ctx = mdt.Context(-1, 1)
    with pytest.raises(mdt.MdtError) as e:
        mdt.Fix(ctx, bad)
    assert e.value.code == code


def test_chemical_options_parse(mdt):
    """STRU_FACTOR ... Chemical and WEIGHT_CHEM take Chem_ctarget, Chem_sigma and Chem_map
    pairs written "(t,w)", "(t w)" or as the two words "(t" "w)" (CVs_Strufactor.cu:125-183)."""
    ctx = mdt.Context(-1, 1)
    tail = " CV_dim 1 SYMBOL v1 S.AVE DIM 1 0 10 100 v1"
    for spec in ("STRU_FACTOR all q_factor 2 Chemical Chem_ctarget 1 Chem_sigma 0.5",
                 "STRU_FACTOR all q_factor 2 Chemical Chem_map (1,1.0) (2,-0.5)",
                 "WEIGHT_CHEM all cutoff_r 5 Chem_map (1 1.0) (2 -0.5)",
                 "WEIGHT_CHEM all Chem_map (1 1.0) (2 -0.5) SW_func FERMI r_0 3"):
        import shlex
        words = ["GAUSSIAN", "0.01", "0.005", "10", "CAL", "NAME", "S"] + shlex.split(spec.replace("(1 1.0)", '"(1 1.0)"')) + tail.split()
        mdt.Fix(ctx, words).close()
    mdt.Fix(ctx, ["GAUSSIAN", "0.01", "0.005", "10", "CAL", "NAME", "S", "WEIGHT_CHEM", "all", "Chem_map", "(1", "2.0)",
                  "(3", "0.5)"] + tail.split()).close()
    ctx.close()


def test_fix_semantic_errors(mdt):
    ctx = mdt.Context(-1, 1)
    base = "GAUSSIAN 0.01 0.005 10 CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
    for tail, frag in [
        ("CV_dim 1 SYMBOL v1 Q7.AVE DIM 1 0 1 100 v1", "not defined"),
        ("CV_dim 1 SYMBOL v1 Q6.MEAN DIM 1 0 1 100 v1", "find the func"),
        ("CV_dim 1 SYMBOL v1 Q6.AVE", "DIM 1 is missing"),
        ("CV_dim 2 SYMBOL v1 Q6.AVE DIM 1 0 1 100 v1", "DIM 2 is missing"),
        ("CV_dim 1 SYMBOL v1 Q6.AVE DIM 2 0 1 100 v1", "out of range"),
        ("CV_dim 3 SYMBOL v1 Q6.AVE DIM 1 0 1 100 v1 DIM 2 0 1 100 v1 DIM 3 0 1 100 v1", "grid supports"),
        ("CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 3 v1 Gaussian_Hill_type 1", "bins must be >= 4"),
        ("CV_dim 3 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 DIM 2 0 1 1000 v1 DIM 3 0 1 1000 v1 Gaussian_Hill_type 1",
         "exceeds"),
        ("CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 100 v1 Gaussian_Hill_type 5", "unknown"),
        ("CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 100 v1 SIGMAS 0.02", "isotropic"),
        ("CAL NAME Q6 STEINH Q 4 all CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 100 v1", "duplicated"),
    ]:
        with pytest.raises(mdt.MdtError) as e:
            mdt.Fix(ctx, base + tail)
        assert frag in str(e.value), (tail, str(e.value))


def test_sparse_hash_grid_parses_in_3d(mdt):
    """Gaussian_Hill_type 1 (the reference's sparse-hash grid, GH_t1_sparseHash.cu) is
    accepted up to three dimensions; the uniform grid (type 0) stays at two."""
    ctx = mdt.Context(-1, 1)
    tail = ("CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 CAL NAME S SK all FCC111 5.26 "
            "CV_dim 3 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE SYMBOL v3 S.AVE "
            "DIM 1 0 1 100 v1 DIM 2 0 1 100 v2 DIM 3 0 100 100 v3 ")
    mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 " + tail + "Gaussian_Hill_type 1").close()
    with pytest.raises(mdt.MdtError) as e:
        mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 " + tail + "Gaussian_Hill_type 0")
    assert "grid supports" in str(e.value)


def test_learned_bias_options(mdt):
    """Gaussian_Hill_type 3 with its MLP_* options parses; out-of-range values and a
    learned bias on fewer than three dimensions are refused."""
    ctx = mdt.Context(-1, 1)
    base = ("GAUSSIAN 0.01 0.005 10 SIGMAS 0.01 0.01 0.5 "
            "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME S SK all FCC111 5.26 "
            "CV_dim 3 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE SYMBOL v3 S.AVE "
            "DIM 1 0 1 100 v1 DIM 2 0 1 100 v2 DIM 3 0 100 100 v3 Gaussian_Hill_type 3 ")
    mdt.Fix(ctx, base + "MLP_REFIT 250 MLP_POINTS 65536 MLP_HOLDOUT 4096 MLP_BATCH 1024 MLP_EPOCHS 4 "
                        "MLP_EPOCHS_FIRST 20 MLP_LR 0.005 MLP_LR_REFIT 0.0005 MLP_TOL 0.1 MLP_SPREAD 2 MLP_SEED 7").close()
    for bad in ("MLP_REFIT 0", "MLP_BATCH 8", "MLP_LR -1", "MLP_LR_REFIT -0.1", "MLP_TOL -1", "MLP_FOO 1"):
        with pytest.raises(mdt.MdtError) as e:
            mdt.Fix(ctx, base + bad)
        assert e.value.code == -22, bad
    with pytest.raises(mdt.MdtError):
        mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
                     "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 100 v1 Gaussian_Hill_type 3")


def test_parse_only_context_refuses_steps(mdt):
    ctx = mdt.Context(-1, 1)
    fix = mdt.Fix(ctx, GOOD)
    with pytest.raises(mdt.MdtError) as e:
        fix.init()
    assert "parse-only" in str(e.value)


@pytest.mark.parametrize("expr,frag", [("v1+v9", "unknown symbol"), ("v1+", "unexpected end"),
                                       ("(v1", "missing ')'"), ("foo(v1)", "unknown function"),
                                       ("v1 v2", "trailing")])
def test_expression_errors(mdt, expr, frag):
    ctx = mdt.Context(-1, 1)
    with pytest.raises(mdt.MdtError) as e:
        mdt.Fix(ctx, GOOD.replace('"v1+v2"', f'"{expr}"'))
    assert frag in str(e.value)


def test_fes_input_parser(mdt):
    """The FES tool takes its grid from the input's fix metad command, joined and stripped by
    LAMMPS's input rules (trailing '&' continues, '#' comments) and read by the library's own
    fix parser on a parse-only context: DIM bounds per index, GAUSSIAN, PACE, METAD_RESTART.
    A keyword list the fix refuses is refused with the parser's message."""
    import os
    import tempfile
    from mdt import fes
    deck = ("units metal\n"
            "# fix 2 all metad GAUSSIAN 9 9 9 (a comment, not a command)\n"
            "fix 1 all metad GAUSSIAN 0.02 0.004 8 PACE 5 &\n"
            "   CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 &\n"
            "   CAL NAME Q4 STEINH Q 4 all CV_dim 2 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE &\n"
            "   DIM 1 0 1 200 v1 DIM 2 -0.5 0.5 150 v2 WT 1 METAD_RESTART 1   # trailing comment\n"
            "run 100\n")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "run.in")
        open(path, "w").write(deck)
        assert fes.fix_words(path)[:6] == ["GAUSSIAN", "0.02", "0.004", "8", "PACE", "5"]
        assert fes.fix_words(path)[-1] == "1"
        p = fes.parse_metad_params(path)
        assert fes.parse_metad_params(os.path.join(tmp, "missing.in")) is None
        open(path, "w").write("units metal\nrun 10\n")
        assert fes.parse_metad_params(path) is None
        open(path, "w").write("fix 1 all metad GAUSSIAN 0.02 0.004 8 CAL NAME Q6 STEINH Q 6 all &\n"
                              "  CV_dim 1 SYMBOL v1 Q6.AVE\n")
        with pytest.raises(mdt.MdtError) as e:
            fes.parse_metad_params(path)
        assert "DIM 1 is missing" in str(e.value)
    assert p["sigma"] == 0.02 and p["height0"] == 0.004 and p["biasf"] == 8.0 and p["pace"] == 5
    assert p["cv_dim"] == 2 and p["continue"] is True
    assert p["bounds"] == {1: [0.0, 1.0, 200], 2: [-0.5, 0.5, 150]}
    axes, lo, d = fes.grid_axes(p)
    assert np.array_equal(axes[0], np.linspace(0.0025, 0.9975, 200))
    assert np.allclose(d, [0.005, 1.0 / 150]) and np.array_equal(lo, [0.0, -0.5])


def test_steinh_l_parsing(mdt):
    """STEINH L with named switching functions: the functions of the CAL (MEAN_SOLID/AVE,
    NSOLID/SW_FUNC, FRAC), one function per CAL, unknown names refused; a SW_FUNC name may
    equal a CV name (separate registries, fix_crystallize.cu:144-160)."""
    ctx = mdt.Context(-1, 1)
    head = ("GAUSSIAN 0.01 0.005 10 CAL NAME sr SW_FUNC RATIONAL r_0 3.0 CAL NAME sc SW_FUNC FERMI r_0 0.4 alpha 10 "
            "CAL NAME L6 STEINH L 6 all cutoff_eps 1e-5 SW_FUNC_r sr SW_FUNC_cv sc ")
    for fn in ("MEAN_SOLID", "AVE", "NSOLID", "SW_FUNC", "FRAC"):
        mdt.Fix(ctx, head + f"CV_dim 1 SYMBOL v1 L6.{fn} DIM 1 0 1 100 v1").close()
    mdt.Fix(ctx, head + "CV_dim 1 SYMBOL v1 L6.AVE SYMBOL v2 L6.MEAN_SOLID DIM 1 0 1 100 v1+v2").close()
    mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 CAL NAME L6 SW_FUNC STEP r_0 3.9 CAL NAME L6 STEINH L 6 all SW_FUNC_r L6 "
                 "CV_dim 1 SYMBOL v1 L6.FRAC DIM 1 0 1 100 v1").close()
    for tail, frag in [
        ("CV_dim 1 SYMBOL v1 L6.COUNT DIM 1 0 1 100 v1", "Use MEAN_SOLID"),
        ("CV_dim 1 SYMBOL v1 L6.NSOLID SYMBOL v2 L6.FRAC DIM 1 0 1 100 v1+v2", "one CAL per function"),
    ]:
        with pytest.raises(mdt.MdtError) as e:
            mdt.Fix(ctx, head + tail)
        assert frag in str(e.value), (tail, str(e.value))
    with pytest.raises(mdt.MdtError) as e:
        mdt.Fix(ctx, "GAUSSIAN 0.01 0.005 10 CAL NAME L6 STEINH L 6 all SW_FUNC_cv nope CV_dim 1 SYMBOL v1 L6.AVE DIM 1 0 1 100 v1")
    assert "not found in registry" in str(e.value)
    ctx.close()
