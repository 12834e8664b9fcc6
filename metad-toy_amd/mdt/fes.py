"""Free-energy surface from a HILLS file, computed on the GPU.

Mirrors the reference's post-processing script tools/plot_hills.py (usage
``python -m mdt.fes <LAMMPS_Input_File> <HILLS_File>``): the grid comes from
the ``fix metad`` line of the LAMMPS input (parse_metad_params, :41-119), the
hills are summed on it (:158-223) and written to FES.dat (:226); for two CVs
the marginal free energies along each CV are written to
Marginal_FES_CV1.dat and Marginal_FES_CV2.dat (:236-285).

The sums run in the library's kernels (mdt_fes_sum, mdt_fes_marginal); this
module only parses, loads and writes.  Two departures from the script: the
HILLS file named on the command line is the one read (the script reads
./HILLS whatever it is given, :8), and no plots are drawn.
"""
import os
import sys

import numpy as np

from . import Context

T = 300.0
R = 8.3144621
KBT = R * T / 1000.0        # the script's k_B T (kJ/mol), used by the marginals


def parse_metad_params(input_file_path):
    """The script's keyword scan of the `fix ... metad` line (tools/plot_hills.py:41-119):
    the tokens after `fix ID group metad`, and every token of the lines after it; CAL
    blocks are skipped up to the next top-level keyword.  Returns None without a fix."""
    try:
        with open(input_file_path) as f:
            content = f.read()
    except FileNotFoundError:
        return None
    content = content.replace("&\n", " ").replace("&", " ")
    tokens, found = [], False
    for line in content.split("\n"):
        line = line.split("#")[0].strip()
        if not line:
            continue
        t = line.split()
        if t[0] == "fix" and "metad" in t:
            tokens.extend(t[4:])
            found = True
        elif found:
            tokens.extend(t)
    if not found:
        return None
    p = {"sigma": 0.05, "height0": 0.1, "biasf": 10.0, "pace": 100, "cv_dim": 1, "nbin_num": 100,
         "continue": False, "bounds": {}}
    stop = ("GAUSSIAN", "PACE", "CV_dim", "DIM", "METAD_RESTART", "WT", "CAL")
    i = 0
    while i < len(tokens):
        a = tokens[i]
        if a == "GAUSSIAN":
            p["sigma"], p["height0"], p["biasf"] = float(tokens[i + 1]), float(tokens[i + 2]), float(tokens[i + 3])
            i += 4
        elif a == "PACE":
            p["pace"] = int(tokens[i + 1])
            i += 2
        elif a == "CV_dim":
            p["cv_dim"] = int(tokens[i + 1])
            i += 2
        elif a == "CAL":
            i += 1
            while i < len(tokens) and tokens[i] not in stop:
                i += 1
        elif a == "DIM":
            try:
                k = int(tokens[i + 1])
                p["bounds"][k] = [float(tokens[i + 2]), float(tokens[i + 3]), int(tokens[i + 4])]
                p["nbin_num"] = int(tokens[i + 4])
            except (ValueError, IndexError):
                pass
            i += 6
        elif a == "METAD_RESTART":
            p["continue"] = int(tokens[i + 1]) != 0
            i += 2
        elif a == "WT":
            i += 2
        else:
            i += 1
    return p


def grid_axes(params):
    """Bin-centre coordinates of each CV as the script builds them (np.linspace from
    lo + d/2 to hi - d/2, :125-156), with the bounds and widths per CV."""
    axes, lo, d = [], [], []
    for k in range(1, params["cv_dim"] + 1):
        a, b, n = params["bounds"].get(k, [0.0, 1.0, 100])
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
