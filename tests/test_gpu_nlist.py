"""Device neighbour-list build (mdt_neighbor_build) against the host generator (run with -m gpu).

Bar: bit-exact.  Ghost positions, ghost owners, list offsets and list entries, in order, equal
the host's (the N-nearest insertion breaks ties by list order, so order is part of the data).
"""
import time

import numpy as np
import pytest

from helpers import fcc_system, liquid_system

pytestmark = pytest.mark.gpu


def _host_and_device(mdt, system, rg):
    x, off, idx, own = system.arrays()
    n = system.n_local
    lo, hi = system.box()
    ctx = mdt.Context(0, 1)
    t0 = time.perf_counter()
    na, nz = ctx.build_neighbors(x[:n], lo, hi, rg)
    t = time.perf_counter() - t0
    xd, offd, idxd, ownd = ctx.fetch_neighbors()
    return ctx, (x, off, idx, own), (xd, offd, idxd, ownd), (na, nz, t)


@pytest.mark.parametrize("case", ["fcc", "liquid", "half_melted", "slab_box"])
def test_device_lists_equal_host_lists(mdt, case):
    rg = 5.9
    if case == "fcc":
        s = fcc_system(mdt, 6, sigma=0.1, seed=71)
    elif case == "liquid":
        s = liquid_system(mdt, 3000, seed=72)
    elif case == "half_melted":
        s = fcc_system(mdt, 12, sigma=0.6, sigma_above=0.1, seed=73)
    else:
        # a box that is not cubic, atoms on its faces included
        s = mdt.System().set_box([0, 0, 0], [30.0, 24.0, 41.0]).add_uniform(2500, 0.0, 41.0, seed=74)
        s.set_positions(np.vstack([s.arrays()[0][:2496], [[0.0, 0.0, 0.0], [29.999999, 0.0, 20.0],
                                                          [15.0, 23.9999999, 0.0], [0.0, 12.0, 40.99999]]]))
        s.build(rg)
    ctx, host, dev, (na, nz, t) = _host_and_device(mdt, s, rg)
    print(f"[nlist] {case}: n={s.n_local} n_all={na} nnz={nz} device build {1e3 * t:.2f} ms", flush=True)
    assert na == s.n_all and nz == host[2].size
    assert np.array_equal(dev[0], host[0])
    assert np.array_equal(dev[1], host[1])
    assert np.array_equal(dev[2], host[2])
    assert np.array_equal(dev[3], host[3])
    ctx.close()


def test_step_on_device_lists_equals_host_lists(mdt):
    """A Q6+Q4 step on lists bound from the device build gives the same bits as on the host's."""
    s = fcc_system(mdt, 7, sigma=0.3, sigma_above=0.1, seed=75)
    n = s.n_local
    args = ("GAUSSIAN 0.01 0.005 10 PACE 1 HILLS_FILE NONE "
            "CAL NAME Q6 STEINH Q 6 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CAL NAME Q4 STEINH Q 4 all cutoff_r 3.9 cutoff_Natoms 18 "
            "CV_dim 1 SYMBOL v1 Q6.AVE SYMBOL v2 Q4.AVE DIM 1 0 2 2000 v1+v2 WT 1")
    out = []
    for use_device in (False, True):
        ctx = mdt.Context(0, 1)
        if use_device:
            lo, hi = s.box()
            ctx.build_neighbors(s.arrays()[0][:n], lo, hi, 5.9)
            ctx.bind_built_neighbors(0)
        else:
            ctx.bind_system(0, s)
        fix = mdt.Fix(ctx, args)
        fix.init()
        f = np.zeros((n, 3))
        for step in range(3):
            fix.post_force(step, f)
        out.append((f.copy(), fix.cv("Q6.AVE"), fix.cv("Q4.AVE"), fix.scalar()))
        fix.close()
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


def test_device_build_at_256k_atoms(mdt):
    """The 256,000-atom half-melted system of the benchmark: the device build matches the host's
    list sizes and a sample of its lists, and reports its time."""
    s = fcc_system(mdt, 40, sigma=0.6, sigma_above=0.1, seed=2024)
    ctx, host, dev, (na, nz, t) = _host_and_device(mdt, s, 5.9)
    lo, hi = s.box()
    t0 = time.perf_counter()
    for _ in range(5):
        ctx.build_neighbors(host[0][: s.n_local], lo, hi, 5.9)
    t5 = (time.perf_counter() - t0) / 5
    print(f"[nlist] 256k: n_all={na} nnz={nz} device build {1e3 * t5:.2f} ms per call (first {1e3 * t:.1f} ms)",
          flush=True)
    assert na == s.n_all and nz == host[2].size
    assert np.array_equal(dev[1], host[1]) and np.array_equal(dev[2], host[2]) and np.array_equal(dev[3], host[3])
    ctx.close()


def test_device_build_of_an_empty_box(mdt):
    """No local atoms: no ghosts, offsets [0], no entries."""
    ctx = mdt.Context(0, 1)
    na, nz = ctx.build_neighbors(np.zeros((0, 3)), [0, 0, 0], [20.0, 20.0, 20.0], 5.9)
    assert (na, nz) == (0, 0)
    x, off, idx, own = ctx.fetch_neighbors()
    assert x.shape == (0, 3) and off.tolist() == [0] and idx.size == 0 and own.size == 0
    with pytest.raises(mdt.MdtError):
        ctx.build_neighbors(np.zeros((4, 3)), [0, 0, 0], [10.0, 20.0, 20.0], 5.9)   # edge < 2 rg
    ctx.close()
