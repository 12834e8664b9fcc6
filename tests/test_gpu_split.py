"""Split reductions (run on an MI355X with -m gpu): two contexts each hold the whole system
but add only their share of the atoms (even / odd) to the CV sums, as two MPI ranks add their
local atoms.  Their partial sums (mdt_fix_partials), summed on the host as MPI_Allreduce
would sum them, resume the step (mdt_fix_resume) in both; the two contexts' forces add up to
those of one context that holds everything, step after step through hill deposits.

Bars: CV values 1e-12 relative (FP64 sums of the same per-atom values, grouped differently;
1e-6 for a local Q6 whose per-atom values come from another pass);
forces 1e-5 relative (FP32 pair sums split in two); hill records 1e-12 relative.
"""
import numpy as np
import pytest

from helpers import fcc_system, rel_force_err

pytestmark = pytest.mark.gpu

Q = "cutoff_r 3.9 cutoff_Natoms 18 "
DECKS = {
    "q6q4_lq6": ("GAUSSIAN 0.01 0.005 10 PACE 2 TEMP 300 HILLS_FILE NONE "
                 f"CAL NAME Q6 STEINH Q 6 all {Q}CAL NAME Q4 STEINH Q 4 all {Q}CAL NAME LQ6 STEINH LD 6 all {Q}"
                 "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE SYMBOL v3 LQ6.AVE DIM 1 0 2 2000 v1+v2 WT 1",
                 ["Q6.AVE", "Q4.AVE", "LQ6.AVE"]),
    "lq6_sk": ("GAUSSIAN 0.01 0.005 10 PACE 2 TEMP 300 HILLS_FILE NONE "
               f"CAL NAME Q6 STEINH Q 6 all {Q}CAL NAME LQ6 STEINH LD 6 all {Q}"
               "CAL NAME SK SK all KVEC 1.1 0.3 -0.2 KVEC -0.4 0.9 0.7 KVEC 0.2 -0.6 1.3 "
               "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 LQ6.AVE SYMBOL v3 SK.AVE DIM 1 0 6 6000 v1+v2+v3 WT 1",
               ["Q6.AVE", "LQ6.AVE", "SK.AVE"]),
}


@pytest.mark.parametrize("deck", sorted(DECKS))
def test_two_shares_add_up_to_the_whole(mdt, deck):
    args, cvs = DECKS[deck]
    s = fcc_system(mdt, 8, sigma=0.3, seed=51)
    n = s.n_local
    print(f"[split] deck={deck} n={n}", flush=True)
    ctx1 = mdt.Context(0, 1)
    ctx1.bind_system(0, s)
    whole = mdt.Fix(ctx1, args)
    parts = []
    for r in (0, 1):
        c = mdt.Context(0, 1)
        c.bind_system(0, s)
        c.bind_share(0, (np.arange(n) % 2 == r).astype(np.uint8))
        parts.append((c, mdt.Fix(c, args)))
    for step in range(5):
        f1 = np.zeros((n, 3))
        whole.post_force(step, f1)
        red = sum(fx.partials(step) for _, fx in parts)
        fs = []
        for _, fx in parts:
            f = np.zeros((n, 3))
            fx.resume(step, red, f)
            fs.append(f)
        for name in cvs:
            v1 = whole.cv(name)
            # an unbiased local Q6 is summed inside the force pass of a whole step, by its own
            # pass in a split one: the same per-atom average in another FP32 order
            tol = 1e-6 if (name == "LQ6.AVE" and deck == "q6q4_lq6") else 1e-12
            for _, fx in parts:
                assert abs(fx.cv(name) - v1) <= tol * abs(v1), (step, name, fx.cv(name), v1)
        assert np.abs(f1).max() > 0.0 or step == 0
        if np.abs(f1).max() > 0.0:
            e = rel_force_err(fs[0] + fs[1], f1)
            assert e <= 1e-5, (step, e)
        for _, fx in parts:
            assert abs(fx.scalar() - whole.scalar()) <= 1e-12 * max(abs(whole.scalar()), 1e-300)
    h1 = whole.hills(1)
    for _, fx in parts:
        h = fx.hills(1)
        assert h.shape == h1.shape and np.abs(h - h1).max() <= 1e-12 * np.abs(h1).max()
    for c, fx in parts:
        with pytest.raises(mdt.MdtError):
            fx.post_force(9)             # a split fix steps only through partials/resume
        fx.close()
        c.close()
    whole.close()
    ctx1.close()


def _cluster_lists(x, rg):
    """Full lists of a finite cluster by brute force, neighbours in ascending index."""
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    nb = [np.nonzero(d2[i] <= rg * rg)[0] for i in range(x.shape[0])]
    off = np.concatenate([[0], np.cumsum([len(v) for v in nb])]).astype(np.int32)
    return nb, off, np.concatenate(nb).astype(np.int32)


def test_domain_decomposed_ranks_add_up(mdt):
    """The multi-rank binding as a domain-decomposed MD code would use it: each rank's context
    holds only its slab of a cluster plus the atoms within two list hops of it (the ghost shell
    of a local-Q CV), as context atoms with their lists cut to the context, and marks its own
    slab as its share.  Partials summed over the ranks resume the step in each; CV values, the
    bias and the forces (each rank's forces scattered back to the atoms they belong to, ghosts
    included) equal those of one context holding the whole cluster, with every CV biased
    (Q6 + Q4 + the Lechner-Dellago LQ6, whose gradient reaches two hops)."""
    rng = np.random.default_rng(71)
    a, m = 5.26, 6
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(m) for j in range(m) for k in range(m)], float)
    x = ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3) * a + rng.normal(scale=0.25, size=(4 * m ** 3, 3)))
    n = x.shape[0]
    rg = 5.9
    nb, off, idx = _cluster_lists(x, rg)
    args = ("GAUSSIAN 0.01 0.005 10 PACE 2 TEMP 300 HILLS_FILE NONE "
            f"CAL NAME Q6 STEINH Q 6 all {Q}CAL NAME Q4 STEINH Q 4 all {Q}CAL NAME LQ6 STEINH LD 6 all {Q}"
            "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE SYMBOL v3 LQ6.AVE DIM 1 0 3 3000 v1+v2+v3 WT 1")
    lo, hi = [-1000.0] * 3, [1000.0] * 3

    def context(xs, offs, idxs, share=None):
        c = mdt.Context(0, 1)
        c.set_box(0, lo, hi)
        c.bind_positions(0, xs, xs.shape[0])
        c.bind_neighbors(0, offs, idxs, np.zeros(0, np.int32))
        if share is not None:
            c.bind_share(0, share)
        return c, mdt.Fix(c, args)

    ctx_w, whole = context(x, off, idx)
    ranks = []
    mid = np.median(x[:, 0])
    for r in (0, 1):
        mine = np.nonzero((x[:, 0] < mid) == (r == 0))[0]
        hop1 = np.unique(np.concatenate([nb[i] for i in mine]))
        hop2 = np.unique(np.concatenate([nb[i] for i in hop1]))
        atoms = np.union1d(np.union1d(mine, hop1), hop2)          # ascending global index
        pos = {g: k for k, g in enumerate(atoms)}
        sub = [[pos[j] for j in nb[g] if j in pos] for g in atoms]
        soff = np.concatenate([[0], np.cumsum([len(v) for v in sub])]).astype(np.int32)
        sidx = np.array([j for v in sub for j in v], np.int32)
        share = np.isin(atoms, mine).astype(np.uint8)
        print(f"[ranks] rank {r}: share {mine.size}, context atoms {atoms.size}", flush=True)
        assert atoms.size < n                                      # a rank holds part of the cluster
        c, fx = context(x[atoms], soff, sidx, share)
        ranks.append((c, fx, atoms))
    for step in range(4):
        fw = np.zeros((n, 3))
        whole.post_force(step, fw)
        red = sum(fx.partials(step) for _, fx, _ in ranks)
        ft = np.zeros((n, 3))
        for _, fx, atoms in ranks:
            f = np.zeros((atoms.size, 3))
            fx.resume(step, red, f)
            np.add.at(ft, atoms, f)
        for name in ("Q6.AVE", "Q4.AVE", "LQ6.AVE"):
            for _, fx, _ in ranks:
                assert abs(fx.cv(name) - whole.cv(name)) <= 1e-12 * abs(whole.cv(name)), (step, name)
        for _, fx, _ in ranks:
            assert abs(fx.scalar() - whole.scalar()) <= 1e-12 * max(abs(whole.scalar()), 1e-300)
            assert fx.cv_weight("LQ6.AVE") == whole.cv_weight("LQ6.AVE")
        if np.abs(fw).max() > 0.0:
            e = rel_force_err(ft, fw)
            print(f"[ranks] step {step}: force rel err {e:.3e}", flush=True)
            assert e <= 1e-5, (step, e)
    assert np.abs(fw).max() > 0.0
    for c, fx, _ in ranks:
        fx.close()
        c.close()
    whole.close()
    ctx_w.close()
