"""Shared builders for the tests: seeded systems, ghost-consistent moves, fix keyword lists."""
import numpy as np


def fcc_system(mdt, n, a=5.26, sigma=0.1, sigma_above=None, seed=1, rg=5.9):
    s = mdt.System().add_fcc(n, n, n, a, sigma, sigma if sigma_above is None else sigma_above,
                             n * a / 2.0, seed=seed)
    s.build(rg)
    return s


def liquid_system(mdt, n_atoms, density=0.021, seed=3, rg=5.9):
    L = (n_atoms / density) ** (1.0 / 3.0)
    s = mdt.System().set_box([0, 0, 0], [L, L, L]).add_uniform(n_atoms, 0.0, L, seed=seed)
    s.build(rg)
    return s


class Mover:
    """Moves local atoms and their ghost images together (ghost = owner + fixed image shift)."""

    def __init__(self, x_all, n_local, owner):
        self.n = n_local
        self.owner = owner.astype(np.int64)
        self.shift = x_all[n_local:] - x_all[self.owner]
        self.x = x_all.copy()

    def set_local(self, xl):
        self.x[: self.n] = xl
        self.x[self.n:] = xl[self.owner] + self.shift
        return self.x


def q46_fix(extra="", pace=10, wt=0, lo=0.0, hi=2.0, nb=2000, expr="v1+v2", rc=3.9, N=18):
    return (f"GAUSSIAN 0.01 0.005 10 PACE {pace} HILLS_FILE NONE "
            f"CAL NAME Q6 STEINH Q 6 all cutoff_r {rc} cutoff_Natoms {N} "
            f"CAL NAME Q4 STEINH Q 4 all cutoff_r {rc} cutoff_Natoms {N} "
            f"CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE DIM 1 {lo} {hi} {nb} {expr} WT {wt} {extra}")


def rel_force_err(F, F_ref):
    return float(np.abs(F - F_ref).max() / max(np.abs(F_ref).max(), 1e-300))
