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
