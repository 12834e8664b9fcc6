"""Free-energy surface from a HILLS file, computed on the GPU.

Mirrors the reference's post-processing script tools/plot_hills.py (usage
``python -m mdt.fes <LAMMPS_Input_File> <HILLS_File>``): the grid comes from
the ``fix metad`` command of the LAMMPS input, read by the library's own
fix parser (the script scans the keywords itself, :41-119), the
hills are summed on it (:158-223) and written to FES.dat (:226); for two CVs
the marginal free energies along each CV are written to
Marginal_FES_CV1.dat and Marginal_FES_CV2.dat (:236-285).

The sums run in the library's kernels (mdt_fes_sum, mdt_fes_marginal); this
module only parses, loads and writes.  Two departures from the script: the
HILLS file named on the command line is the one read (the script reads
./HILLS whatever it is given, :8), and no plots are drawn.
"""
import os
import shlex
import sys

import numpy as np

from . import Context, Fix

T = 300.0
R = 8.3144621
KBT = R * T / 1000.0        # the script's k_B T (kJ/mol), used by the marginals


def fix_words(input_file_path):
    """Words of the first ``fix <ID> <group> metad ...`` command of a LAMMPS input file,
    read by LAMMPS's own input rules: a line whose last non-blank character is ``&``
    continues on the next line, ``#`` outside quotes starts a comment, quoted words keep
    their blanks.  Returns the words after ``metad``, or None (no file, no such fix)."""
    try:
        with open(input_file_path) as f:
            raw = f.read().splitlines()
    except FileNotFoundError:
        return None
    command = ""
    for line in raw:
        body = line.rstrip()
        if body.endswith("&"):
            command += body[:-1] + " "
            continue
        command += body
        words = shlex.split(command, comments=True)
        command = ""
        if len(words) >= 4 and words[0] == "fix" and words[3] == "metad":
            return words[4:]
    return None


def parse_metad_params(input_file_path):
    """Grid and hill parameters of the input's ``fix metad``, read by the library's own
    keyword parser on a parse-only context (mdt_fix_param): the values the running fix
    would use.  Returns None when the file or the fix is missing; a keyword list the fix
    itself would refuse raises MdtError with the parser's message."""
    words = fix_words(input_file_path)
    if words is None:
        return None
    ctx = Context(-1, 1)
    try:
        fx = Fix(ctx, words)
        try:
            sigma, height0, biasf = (float(v) for v in fx.param("GAUSSIAN"))
            D = int(fx.param("CV_dim")[0])
            dims = fx.param("DIM").reshape(D, 3)
            p = {"sigma": sigma, "height0": height0, "biasf": biasf, "pace": int(fx.param("PACE")[0]),
                 "cv_dim": D, "continue": bool(fx.param("METAD_RESTART")[0]),
                 "bounds": {k + 1: [float(a), float(b), int(n)] for k, (a, b, n) in enumerate(dims)}}
            p["nbin_num"] = p["bounds"][D][2]
        finally:
            fx.close()
    finally:
        ctx.close()
    return p


def grid_axes(params):
    """Bin-centre coordinates of each CV as the script builds them (np.linspace from
    lo + d/2 to hi - d/2, :125-156), with the bounds and widths per CV."""
    axes, lo, d = [], [], []
    for k in range(1, params["cv_dim"] + 1):
        a, b, n = params["bounds"][k]
        w = (b - a) / n
        axes.append(np.linspace(a + w / 2, b - w / 2, n))
        lo.append(a)
        d.append(w)
    return axes, np.array(lo), np.array(d)


def compute(ctx, params, hills):
    """(axes, F) on the GPU; hills rows (step, s..., height, sigma)."""
    D = params["cv_dim"]
    if D not in (1, 2):
        raise ValueError("the FES reconstruction covers CV_dim 1 and 2, as the reference script does")
    hills = np.atleast_2d(np.asarray(hills, np.float64))
    if hills.shape[1] != D + 3:
        raise ValueError(f"HILLS rows have {hills.shape[1]} columns; CV_dim {D} needs {D + 3} "
                         "(step, cv..., height, sigma)")
    axes, lo, d = grid_axes(params)
    from . import fes_sum
    if D == 1:
        return axes, fes_sum(ctx, hills, axes[0])
    return axes, fes_sum(ctx, hills, axes[0], axes[1], lo, d)


def main(argv=None, device=0):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 2:
        print("usage: python -m mdt.fes <LAMMPS_Input_File> <HILLS_File>", file=sys.stderr)
        return 1
    params = parse_metad_params(argv[0])
    if params is None:
        print(f"no fix metad line in {argv[0]}", file=sys.stderr)
        return 1
    hills = np.loadtxt(argv[1], comments="#", ndmin=2)
    ctx = Context(device, 1)
    try:
        axes, F = compute(ctx, params, hills)
        np.savetxt("FES.dat", F)
        if params["cv_dim"] == 2:
            from . import fes_marginal
            _, _, d = grid_axes(params)
            f1 = fes_marginal(ctx, F, 1, KBT, d[1])
            np.savetxt("Marginal_FES_CV1.dat", np.stack((axes[0], f1), axis=1), header="CV1 FES_CV1",
                       fmt="%.8e", comments="#")
            f2 = fes_marginal(ctx, F, 0, KBT, d[0])
            np.savetxt("Marginal_FES_CV2.dat", np.stack((axes[1], f2), axis=1), header="CV2 FES_CV2",
                       fmt="%.8e", comments="#")
    finally:
        ctx.close()
    print(f"FES.dat written ({os.getcwd()})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
