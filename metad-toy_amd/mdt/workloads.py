"""Synthetic systems and fix keyword lists of the five benchmark workloads.

The standalone driver's stand-in for a LAMMPS data file and input deck: seeded
FCC lattices and uniform liquids built by the library's host generator
(mdt_system_*), and the `fix metad` keyword lists run on them.  The benchmark
and the full-size parity tests build the same systems from here.

  c1  4,000 Ar atoms, 10x10x10 FCC cells (a = 5.26 A), jitter 0.1 A; Q6.AVE on [0,1] x 1000
  c2  256,000 atoms, 40^3 FCC cells, lower half jittered 0.6 A, upper half 0.1 A;
      Q6.AVE + Q4.AVE + Lechner-Dellago local Q6, DIM "v1+v2" on [0,2] x 2000, WT
  c3  500,000-atom slab: 52x52x26 FCC cells below L/2, uniform liquid above;
      S(k) over the 8 (111) vectors + 2,048 seeded random k, [0,50] x 1000
  c4  64 walkers of 2,000-64,000 atoms (FCC or liquid at 0.021 / A^3), 2-D grids on
      (Q4.AVE, Q6.AVE) over [0,1]^2 x 200^2, WT
  c5  the c2 system with a 3-D CV (Q4.AVE, Q6.AVE, S(k_111)), widths 0.01/0.01/0.5, WT
"""
import numpy as np

A, RC, RG = 5.26, 3.9, 5.9       # lattice constant, Steinhardt cutoff, list cutoff (A)
KB = 8.617343e-5                 # eV/K


def steinh(name, l, kind="Q", N=18):
    return f"CAL NAME {name} STEINH {kind} {l} all cutoff_r {RC} cutoff_Natoms {N} "


C1_ARGS = ("GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE NONE TEMP 300 " + steinh("Q6", 6) +
           "CV_dim 1 SYMBOL v1 Q6.AVE DIM 1 0 1 1000 v1 WT 0")

C2_ARGS = ("GAUSSIAN 0.01 0.005 10 PACE 10 TEMP 300 HILLS_FILE NONE " + steinh("Q6", 6) + steinh("Q4", 4) +
           steinh("LQ6", 6, "LD") + "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE SYMBOL v3 LQ6.AVE "
           "DIM 1 0 2 2000 v1+v2 WT 1")

SK_SPEC = f"FCC111 {A} NRAND 2048 KMIN 1.5 KMAX 3.5 SEED 31"
C3_ARGS = (f"GAUSSIAN 0.5 0.005 10 PACE 10 HILLS_FILE NONE CAL NAME SK SK all {SK_SPEC} "
           "CV_dim 1 SYMBOL v1 SK.AVE DIM 1 0 50 1000 v1 WT 0")

C4_ARGS = ("GAUSSIAN 0.01 0.005 10 PACE 10 TEMP 300 HILLS_FILE NONE " + steinh("Q4", 4) + steinh("Q6", 6) +
           "CV_dim 2 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE DIM 1 0 1 200 v1 DIM 2 0 1 200 v2 WT 1")

LQ_ARGS = ("GAUSSIAN 0.01 0.005 10 PACE 10 HILLS_FILE NONE CAL NAME sr SW_FUNC RATIONAL r_0 3.0 "
           "CAL NAME sc SW_FUNC FERMI r_0 0.35 alpha 10 "
           "CAL NAME LQ STEINH L 6 all SW_FUNC_r sr SW_FUNC_cv sc "
           "CV_dim 1 SYMBOL v1 LQ.MEAN_SOLID DIM 1 0 2 2000 v1 WT 1")

C5_SIGMAS = np.array([0.01, 0.01, 0.5])
C5_BASE = ("GAUSSIAN 0.01 0.005 10 PACE 10 TEMP 300 HILLS_FILE NONE SIGMAS 0.01 0.01 0.5 " + steinh("Q4", 4) +
           steinh("Q6", 6) + f"CAL NAME S111 SK all FCC111 {A} CV_dim 3 SYMBOL v1 Q4.AVE SYMBOL v2 Q6.AVE "
           "SYMBOL v3 S111.AVE DIM 1 0 1 100 v1 DIM 2 0 1 100 v2 DIM 3 0 200000 100 v3 WT 1 ")


def c1_system(mdt):
    return mdt.System().add_fcc(10, 10, 10, A, 0.1, 0.1, 0.0, seed=1).build(RG)


def c2_system(mdt):
    n = 40
    return mdt.System().add_fcc(n, n, n, A, 0.6, 0.1, n * A / 2.0, seed=2024).build(RG)


def c3_system(mdt):
    L = 52 * A
    s = mdt.System().set_box([0, 0, 0], [L, L, L]).add_fcc(52, 52, 26, A, 0.0, 0.0, 0.0, seed=3)
    n_fcc = 4 * 52 * 52 * 26
    return s.add_uniform(500000 - n_fcc, L / 2, L, seed=4).build(RG)


def c4_systems(mdt, nw=64, seed=44):
    rng = np.random.default_rng(seed)
    out = []
    for w in range(nw):
        nt = int(rng.integers(2000, 64001))
        if rng.random() < 0.5:
            m = max(4, int(round((nt / 4) ** (1 / 3))))
            s = mdt.System().add_fcc(m, m, m, A, 0.1, 0.1, 0.0, seed=1000 + w)
        else:
            L = (nt / 0.021) ** (1 / 3)
            s = mdt.System().set_box([0, 0, 0], [L, L, L]).add_uniform(nt, 0.0, L, seed=2000 + w)
        out.append(s.build(RG))
    return out


def wt_hills(s0, nh, kt, biasf=10.0, h0=0.005, seed=5, sigma=C5_SIGMAS):
    """Hills of a well-tempered run diffusing around s0: centres spread 2 sigma, heights
    h0 exp(-V(c)/(kT (biasf-1))) with V the FP64 sum of the earlier hills."""
    sig = np.asarray(sigma, np.float64)
    rng = np.random.default_rng(seed)
    c = s0 + rng.normal(size=(nh, len(sig))) * 2.0 * sig
    h = np.zeros(nh)
    for k in range(nh):
        V = float(np.sum(h[:k] * np.exp(-0.5 * np.sum(((c[:k] - c[k]) / sig) ** 2, axis=1)))) if k else 0.0
        h[k] = h0 * np.exp(-V / (kt * (biasf - 1.0)))
    return c, h
