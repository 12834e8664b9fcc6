"""Pins of the FP64 oracle (CPU only).

Each oracle function is checked against something other than its own formula:
an independent library (scipy), a closed form, a hand-derived known answer,
finite differences or an invariant.
"""
import numpy as np
import pytest
from scipy.special import sph_harm_y

from helpers import fcc_system


# ------------------------------------------------------------------ Y_lm
@pytest.mark.parametrize("l", [3, 4, 6])
def test_ylm_matches_scipy(orc, l):
    rng = np.random.default_rng(100 + l)
    for _ in range(50):
        d = rng.normal(size=3) * rng.uniform(0.5, 4.0)
        r = np.linalg.norm(d)
        th, ph = np.arccos(d[2] / r), np.arctan2(d[1], d[0])
        ref = np.array([sph_harm_y(l, m, th, ph) for m in range(l + 1)])
        assert np.abs(orc.ylm(l, d) - ref).max() < 1e-13
        # the polynomial form c_lm(u) (x + i y)^m used by the kernels is the same function
        assert np.abs(orc.ylm_poly(l, d) - ref).max() < 1e-13


@pytest.mark.parametrize("l", [3, 4, 6])
def test_ylm_sum_rule(orc, l):
    rng = np.random.default_rng(7)
    for _ in range(20):
        Y = orc.ylm(l, rng.normal(size=3))
        tot = abs(Y[0]) ** 2 + 2 * np.sum(np.abs(Y[1:]) ** 2)
        assert abs(tot - (2 * l + 1) / (4 * np.pi)) < 1e-13


@pytest.mark.parametrize("l", [3, 4, 6])
def test_ylm_gradient_fd(orc, l):
    rng = np.random.default_rng(11)
    h = 1e-6
    for _ in range(10):
        d = rng.normal(size=3) * 3.0
        g = orc.ylm_grad(l, d)
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            fd = (orc.ylm(l, d + e) - orc.ylm(l, d - e)) / (2 * h)
            assert np.abs(fd - g[:, a]).max() < 1e-7


# ------------------------------------------------------------------ q_l known answers
def _shell_q(orc, l, vecs):
    N = len(vecs)
    sel_d = np.zeros((1, N, 4))
    sel_d[0, :, :3] = vecs
    sel_d[0, :, 3] = np.sum(np.asarray(vecs) ** 2, axis=1)
    q = np.zeros(2 * (l + 1))
    ql = np.zeros(1)
    orc.lib().orc_steinhardt(l, 1, N, sel_d.ctypes.data, np.array([N], np.int32).ctypes.data,
                             q.ctypes.data, ql.ctypes.data)
    return ql[0]


def _fcc12():
    v = []
    for a in (-1, 1):
        for b in (-1, 1):
            v += [(a, b, 0), (a, 0, b), (0, a, b)]
    return np.array(v, float) / 2


def _bcc8():
    return np.array([(a, b, c) for a in (-1, 1) for b in (-1, 1) for c in (-1, 1)], float) / 2


def _sc6():
    return np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], float)


def _hcp12():
    a = 1.0
    c = np.sqrt(8.0 / 3.0) * a
    v = [(a * np.cos(k * np.pi / 3), a * np.sin(k * np.pi / 3), 0.0) for k in range(6)]
    for z in (c / 2, -c / 2):
        for k in range(3):
            ang = np.pi / 6 + k * 2 * np.pi / 3
            v.append((a / np.sqrt(3) * np.cos(ang), a / np.sqrt(3) * np.sin(ang), z))
    return np.array(v)


@pytest.mark.parametrize("shell,q4,q6", [
    (_fcc12, 0.190941, 0.574524),
    (_bcc8, 0.509175, 0.628539),
    (_sc6, 0.763763, 0.353553),
    (_hcp12, 0.097222, 0.484762),
])
def test_q_known_answers(orc, shell, q4, q6):
    v = shell()
    assert abs(_shell_q(orc, 4, v) - q4) < 1e-6
    assert abs(_shell_q(orc, 6, v) - q6) < 1e-6


def test_perfect_fcc_system(mdt, orc):
    s = fcc_system(mdt, 5, sigma=0.0)
    x, off, idx, own = s.arrays()
    env = orc.Env(x, off, idx, own, s.n_local, 3.9, 18)
    assert np.all(env.nsel == 12)
    assert abs(env.q_ave(6) - 0.574524) < 1e-6 and abs(env.q_ave(4) - 0.190941) < 1e-6
    # local Q of a perfect lattice equals Q (all neighbours identical)
    _, lq = env.ld(6)
    assert np.abs(lq - 0.5745242597140697).max() < 1e-12


# ------------------------------------------------------------------ selection
def _select(orc, r2_list, N, rc=10.0):
    """One center at the origin with neighbours on the x axis at the given r^2, in list order."""
    n = len(r2_list)
    x = np.zeros((n + 1, 3))
    x[1:, 0] = np.sqrt(np.asarray(r2_list, float))
    off = np.array([0, n], np.int32)
    idx = np.arange(1, n + 1, dtype=np.int32)
    sel_idx = np.zeros(N, np.int32)
    sel_pos = np.zeros(N, np.int32)
    sel_d = np.zeros(N * 4)
    cnt = np.zeros(1, np.int32)
    ns = np.zeros(1, np.int32)
    orc.lib().orc_select(1, np.array([0], np.int32).ctypes.data, off.ctypes.data, idx.ctypes.data, x.ctypes.data,
                         rc * rc, N, 1e6, sel_idx.ctypes.data, sel_pos.ctypes.data, sel_d.ctypes.data,
                         cnt.ctypes.data, ns.ctypes.data)
    return sel_pos.tolist(), cnt[0], ns[0]


def test_selection_known_answers(orc):
    # below the cap with distinct r^2: ascending r^2, everything kept
    pos, cnt, ns = _select(orc, [4.0, 1.0, 9.0, 2.25], 6)
    assert pos == [1, 3, 0, 2, -1, -1] and cnt == 4 and ns == 4
    # above the cap, distinct: the N smallest
    pos, cnt, ns = _select(orc, [4.0, 1.0, 9.0, 2.25, 0.25], 3)
    assert pos == [4, 1, 3] and cnt == 5 and ns == 3
    # ties, worked by hand from the strict '>' carry rule of the reference:
    # a(1), b(1), c(0.25): after a,b -> [a,b]; c displaces a, a passes b (1 > 1 false)
    # and lands behind it -> [c, b, a].  The equal run rotates.
    pos, _, _ = _select(orc, [1.0, 1.0, 0.25], 3)
    assert pos == [2, 1, 0]
    # same with a cap of 2: the rotated 'a' falls off the end
    pos, cnt, ns = _select(orc, [1.0, 1.0, 0.25], 2)
    assert pos == [2, 1] and cnt == 3 and ns == 2
    # an entry equal to the current maximum of a full buffer is dropped
    pos, _, _ = _select(orc, [1.0, 2.0, 2.0], 2)
    assert pos == [0, 1]
    # cutoff and self-distance filters
    pos, cnt, _ = _select(orc, [0.0, 1.0, 100.0 + 1e-9], 4)
    assert cnt == 1 and pos[0] == 1


def test_selection_bruteforce_no_ties(orc):
    """Without ties the kept set is the N smallest r^2 in ascending order."""
    rng = np.random.default_rng(3)
    for _ in range(30):
        r2 = rng.uniform(0.5, 150.0, size=rng.integers(1, 40))
        N = int(rng.integers(1, 20))
        pos, cnt, ns = _select(orc, r2, N)
        keep = [k for k in range(len(r2)) if r2[k] <= 100.0]
        order = sorted(keep, key=lambda k: r2[k])[:N]
        assert cnt == len(keep) and ns == min(len(keep), N)
        assert pos[:ns] == order


def test_tie_case_separates_insertion_from_stable_sort(mdt, orc):
    """The GPU selection test's tie case (perfect FCC, N = 8 of 12 equidistant neighbours)
    tells the reference insertion apart from the nearest-N rule with list-order ties: a
    kernel that kept the stable N smallest, the likeliest slip, keeps a different set for
    hundreds of atoms there, so it fails the set comparison of the GPU test."""
    s = fcc_system(mdt, 6, sigma=0.0)
    x, off, idx, own = s.arrays()
    n, N, rc = s.n_local, 8, 3.9
    env = orc.Env(x, off, idx, own, n, rc, N)
    differ = 0
    for i in range(n):
        seg = idx[off[i]:off[i + 1]]
        d = x[seg] - x[i]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        ok = np.nonzero((r2 <= rc * rc) & (r2 >= 1e-12))[0]
        stable = sorted(ok, key=lambda p: (r2[p], p))[:N]
        differ += set(seg[stable]) != set(env.sel_idx[i][env.sel_idx[i] >= 0])
    assert differ > 100


# ------------------------------------------------------------------ gradients
def _move_all(x_all, n, owner, k, a, h):
    y = x_all.copy()
    y[k, a] += h
    gh = np.nonzero(owner == k)[0] + n
    y[gh, a] += h
    return y


@pytest.mark.parametrize("kind", ["Q6", "Q4", "LQ6", "Q3"])
def test_gradient_finite_differences(mdt, orc, kind):
    s = fcc_system(mdt, 3, sigma=0.25, seed=5)
    x, off, idx, own = s.arrays()
    n = s.n_local
    l = int(kind[-1])
    env = orc.Env(x, off, idx, own, n, 3.9, 18)
    g = env.dLQ_dx(l) if kind.startswith("LQ") else env.dQ_dx(l)

    def cv(y):
        e = orc.Env(y, off, idx, own, n, 3.9, 18)
        return e.ld(l)[1].mean() if kind.startswith("LQ") else e.q_ave(l)

    h = 1e-6
    for k in (0, 7, 50, 101):
        for a in range(3):
            fd = (cv(_move_all(x, n, own, k, a, h)) - cv(_move_all(x, n, own, k, a, -h))) / (2 * h)
            assert abs(fd - g[k, a]) <= 1e-6 * np.abs(g).max() + 1e-9
    # translation invariance
    assert np.abs(g.sum(axis=0)).max() < 1e-12 * n


def test_gradient_with_binding_cap(mdt, orc):
    """N smaller than the shell: asymmetric selections, exact gradient still matches FD."""
    s = fcc_system(mdt, 3, sigma=0.25, seed=8)
    x, off, idx, own = s.arrays()
    n = s.n_local
    env = orc.Env(x, off, idx, own, n, 3.9, 6)
    assert (env.cnt > 6).sum() > n // 2
    g = env.dQ_dx(6)
    h = 1e-6
    for k in (3, 40, 77):
        for a in range(3):
            p = orc.Env(_move_all(x, n, own, k, a, h), off, idx, own, n, 3.9, 6)
            m = orc.Env(_move_all(x, n, own, k, a, -h), off, idx, own, n, 3.9, 6)
            if not (np.array_equal(p.sel_idx, env.sel_idx) and np.array_equal(m.sel_idx, env.sel_idx)):
                continue   # selection changed within +-h: the CV is not differentiable there
            fd = (p.q_ave(6) - m.q_ave(6)) / (2 * h)
            assert abs(fd - g[k, a]) <= 1e-6 * np.abs(g).max() + 1e-9


def test_group_gradient_fd(mdt, orc):
    """CAL on a group: Q6.AVE averages over the group's centers only; the gradient of that
    average reaches group and non-group atoms alike and matches central differences."""
    s = fcc_system(mdt, 4, sigma=0.2, seed=3)
    x, off, idx, own = s.arrays()
    n = s.n_local
    grp = (x[:n, 0] < x[:n, 0].mean()).astype(np.uint8)
    env = orc.Env(x, off, idx, own, n, 3.9, 18, group=grp)
    q, ql = env.steinhardt(6)
    assert abs(env.q_ave(6) - ql.sum() / grp.sum()) < 1e-15
    g = env.dQ_dx(6)
    h = 1e-6
    outside = np.nonzero(grp == 0)[0]
    for k in (int(np.nonzero(grp)[0][3]), int(outside[5]), 100):
        for a in range(3):
            p = orc.Env(_move_all(x, n, own, k, a, h), off, idx, own, n, 3.9, 18, group=grp)
            m = orc.Env(_move_all(x, n, own, k, a, -h), off, idx, own, n, 3.9, 18, group=grp)
            fd = (p.q_ave(6) - m.q_ave(6)) / (2 * h)
            assert abs(fd - g[k, a]) <= 1e-6 * np.abs(g).max() + 1e-9


# ------------------------------------------------------------------ grid
def test_grid_single_hill_closed_form(orc):
    g = orc.Grid([0.0], [1.0], [1000])
    g.deposit([0.3217], 0.005, 0.01)
    ctr = (np.arange(1000) + 0.5) / 1000
    full = 0.005 * np.exp(-0.5 * (0.3217 - ctr) ** 2 / 0.01 ** 2)
    R = int(np.ceil(4 * 0.01 / 0.001)) + 1
    assert R == 41
    c = int(0.3217 / 0.001)
    win = np.zeros(1000, bool)
    win[max(0, c - R):min(999, c + R) + 1] = True
    assert np.abs(g.g[win] - full[win]).max() < 1e-15
    assert np.all(g.g[~win] == 0)
    # truncation bound h exp(-(R-1)^2 dx^2 / (2 sigma^2))
    assert full[~win].max() <= 0.005 * np.exp(-((R - 1) * 0.001) ** 2 / (2 * 0.01 ** 2))


def test_catmull_rom_exact_on_quadratics(orc):
    for D in (1, 2):
        lo, hi, nb = [0.0] * D, [2.0] * D, [40] * D
        g = orc.Grid(lo, hi, nb)
        c = (np.arange(40) + 0.5) * 0.05
        if D == 1:
            g.g[:] = 0.3 + 1.7 * c - 0.9 * c ** 2
        else:
            X, Y = np.meshgrid(c, c, indexing="ij")
            g.g[:] = (0.3 + 1.7 * X - 0.9 * X ** 2 + 0.4 * Y - 0.2 * Y ** 2 + 0.5 * X * Y).ravel()
        rng = np.random.default_rng(D)
        for _ in range(20):
            s = rng.uniform(0.2, 1.8, size=D)
            V, dV = g.eval(s)
            if D == 1:
                x0 = s[0]
                assert abs(V - (0.3 + 1.7 * x0 - 0.9 * x0 ** 2)) < 1e-12
                assert abs(dV[0] - (1.7 - 1.8 * x0)) < 1e-10
            else:
                x0, y0 = s
                assert abs(V - (0.3 + 1.7 * x0 - 0.9 * x0 ** 2 + 0.4 * y0 - 0.2 * y0 ** 2 + 0.5 * x0 * y0)) < 1e-12
                assert abs(dV[0] - (1.7 - 1.8 * x0 + 0.5 * y0)) < 1e-10
                assert abs(dV[1] - (0.4 - 0.4 * y0 + 0.5 * x0)) < 1e-10


def test_grid_clamp_and_constant(orc):
    g = orc.Grid([0.0], [1.0], [100])
    g.g[:] = 0.7
    for s in (-1.0, 0.0, 0.004, 0.5, 0.99, 2.0):
        V, dV = g.eval([s])
        assert abs(V - 0.7) < 1e-15 and abs(dV[0]) < 1e-12
    g.g[:] = np.arange(100.0)
    # below lo + 1.5 dx the index clamps to 1 with f = 0: value = g[1]
    assert g.eval([0.0])[0] == 1.0
    # above hi - 2.5 dx it clamps to nb-3 with f = 1: value = g[nb-2]
    assert g.eval([1.0])[0] == 98.0


def test_wt_height(orc):
    g = orc.Grid([0.0], [1.0], [1000])
    s = [0.4]
    assert g.wt_height(s, 0.005, 10.0, 300.0, 8.617343e-5, 1) == 0.005
    g.deposit(s, 0.005, 0.01)
    V = g.eval(s)[0]
    kT = 300.0 * 8.617343e-5
    assert abs(g.wt_height(s, 0.005, 10.0, 300.0, 8.617343e-5, 1) - 0.005 * np.exp(-V / (kT * 9.0))) < 1e-18
    assert g.wt_height(s, 0.005, 10.0, 300.0, 8.617343e-5, 0) == 0.005


def test_sparse_hash_grid_3d_single_hill_closed_form(orc):
    """Sparse-hash deposit (GH_t1_sparseHash.cu:228-262): window of R = ceil(4 sigma/dx) + 1
    bins around c = int((s - lo)/dx) per dimension, exp(-sum dx^2 / (2 sigma^2)) inside, 0 outside."""
    lo, hi, nb = np.array([0.0, 0.0, 10.0]), np.array([1.0, 0.5, 30.0]), np.array([50, 40, 30])
    g = orc.GridT1(lo, hi, nb)
    s, w, sig = np.array([0.4321, 0.2101, 18.77]), 0.005, 0.04
    g.deposit(s, w, sig)
    dx = (hi - lo) / nb
    ctr = [lo[d] + (np.arange(nb[d]) + 0.5) * dx[d] for d in range(3)]
    C = np.meshgrid(*ctr, indexing="ij")
    full = w * np.exp(-0.5 * sum((s[d] - C[d]) ** 2 for d in range(3)) / sig ** 2)
    win = np.ones(tuple(nb), bool)
    for d in range(3):
        R = int(np.ceil(4 * sig / dx[d])) + 1
        c = int((s[d] - lo[d]) / dx[d])
        m = np.zeros(nb[d], bool)
        m[max(0, c - R):min(nb[d] - 1, c + R) + 1] = True
        win &= m.reshape([-1 if k == d else 1 for k in range(3)])
    G = g.g.reshape(tuple(nb))
    assert np.abs(G[win] - full[win]).max() < 1e-15
    assert np.all(G[~win] == 0)


def test_sparse_hash_grid_catmull_rom_exact_on_triquadratics(orc):
    """The tensor Catmull-Rom of the recursion (GH_t1_sparseHash.cu:316-366) reproduces
    polynomials of degree <= 2 in each variable: V and every dV/ds_d exactly."""
    lo, hi, nb = [0.0, -1.0, 2.0], [2.0, 1.0, 5.0], [30, 24, 20]
    g = orc.GridT1(lo, hi, nb)
    ctr = [lo[d] + (np.arange(nb[d]) + 0.5) * (hi[d] - lo[d]) / nb[d] for d in range(3)]
    X, Y, Z = np.meshgrid(*ctr, indexing="ij")

    def P(x, y, z):
        return 0.3 + 1.7 * x - 0.9 * x * x + 0.4 * y * z - 0.2 * z * z + 0.5 * x * y * z + 0.1 * x * x * y * y * z * z

    def dP(x, y, z):
        return np.array([1.7 - 1.8 * x + 0.5 * y * z + 0.2 * x * y * y * z * z,
                         0.4 * z + 0.5 * x * z + 0.2 * x * x * y * z * z,
                         0.4 * y - 0.4 * z + 0.5 * x * y + 0.2 * x * x * y * y * z])
    g.g[:] = P(X, Y, Z).ravel()
    rng = np.random.default_rng(11)
    for _ in range(30):
        sp = np.array([rng.uniform(0.2, 1.8), rng.uniform(-0.8, 0.8), rng.uniform(2.3, 4.6)])
        V, dV = g.eval(sp)
        assert abs(V - P(*sp)) < 1e-11
        assert np.abs(dV - dP(*sp)).max() < 1e-9


def test_sparse_hash_grid_agrees_with_uniform_grid(orc):
    """In one and two dimensions the sparse-hash grid and the uniform grid deposit the same
    Gaussians at the same bins and interpolate the same spline (the evaluation order differs,
    so they agree to rounding)."""
    rng = np.random.default_rng(3)
    for D in (1, 2):
        lo, hi, nb = [0.0] * D, [1.0] * D, [120] * D
        a, b = orc.Grid(lo, hi, nb), orc.GridT1(lo, hi, nb)
        for _ in range(40):
            s = rng.uniform(0.1, 0.9, size=D)
            a.deposit(s, 0.005, 0.03)
            b.deposit(s, 0.005, 0.03)
        assert np.abs(a.g - b.g).max() <= 1e-15 * max(1.0, np.abs(a.g).max()) * 40
        for _ in range(20):
            s = rng.uniform(0.0, 1.0, size=D)
            Va, dVa = a.eval(s)
            Vb, dVb = b.eval(s)
            assert abs(Va - Vb) <= 1e-12 * max(1e-300, abs(Va)) + 1e-18
            assert np.abs(dVa - dVb).max() <= 1e-10 * max(1e-300, np.abs(dVa).max()) + 1e-16


def test_sparse_hash_grid_clamp_and_constant(orc):
    g = orc.GridT1([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [10, 12, 14])
    g.g[:] = 0.7
    for s in ([-1.0, 0.5, 0.5], [0.5, 2.0, 0.5], [0.0, 0.0, 1.0], [0.3, 0.6, 0.9]):
        V, dV = g.eval(s)
        assert abs(V - 0.7) < 1e-15 and np.abs(dV).max() < 1e-12
    g.g[:] = np.arange(10 * 12 * 14, dtype=float)
    # every coordinate below lo + 1.5 dx clamps to bin 1 with f = 0
    assert g.eval([0.0, 0.0, 0.0])[0] == (1 * 12 + 1) * 14 + 1
    # above hi - 2.5 dx to nb - 3 with f = 1: the value of bin nb - 2
    assert g.eval([1.0, 1.0, 1.0])[0] == (8 * 12 + 10) * 14 + 12


def test_reference_hills_fixture(orc):
    """The reference's own HILLS file (238 1-D hills, h 0.005, sigma 0.01) on the config-1 grid."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "HILLS_reference_test_cvset")
    rows = np.loadtxt(path, comments="#")
    assert rows.shape == (238, 4)
    g = orc.Grid([0.0], [1.0], [1000])
    for r in rows:
        g.deposit([r[1]], r[2], r[3])
    ctr = (np.arange(1000) + 0.5) / 1000
    full = np.sum(rows[:, 2, None] * np.exp(-0.5 * (rows[:, 1, None] - ctr) ** 2 / rows[:, 3, None] ** 2), axis=0)
    # truncated windows differ from the full sum by at most 238 tails of h e^-8
    assert np.abs(g.g - full).max() <= 238 * 0.005 * np.exp(-8.0)
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "grid_reference_hills.npy"))
    assert np.array_equal(g.g, gold)


# ------------------------------------------------------------------ hills sum
def test_hills_sum_fd_and_grid(orc):
    rng = np.random.default_rng(4)
    c = rng.uniform(0.3, 0.7, size=(50, 3))
    h = rng.uniform(0.001, 0.005, size=50)
    sg = np.array([0.01, 0.02, 0.5])
    s = rng.uniform(0.3, 0.7, size=(5, 3))
    V, dV = orc.hills_sum(c, h, sg, s)
    for p in range(5):
        for a in range(3):
            e = np.zeros(3)
            e[a] = 1e-6
            fd = (orc.hills_sum(c, h, sg, s[p] + e)[0][0] - orc.hills_sum(c, h, sg, s[p] - e)[0][0]) / 2e-6
            assert abs(fd - dV[p, a]) < 1e-6 * max(1.0, abs(dV[p, a]))
    # 1-D: equals the grid at a bin centre when every hill lies well inside the window
    g = orc.Grid([0.0], [1.0], [1000])
    c1 = np.array([[0.4], [0.41]])
    for k in range(2):
        g.deposit(c1[k], 0.005, 0.01)
    Vs, _ = orc.hills_sum(c1, [0.005, 0.005], [0.01], [[0.4055]])
    assert abs(Vs[0] - g.g[405]) < 1e-15


# ------------------------------------------------------------------ structure factor
def test_sk_lattice_and_symmetry(mdt, orc):
    s = fcc_system(mdt, 4, sigma=0.0)
    x = s.arrays()[0][: s.n_local]
    a = 5.26
    k111 = 2 * np.pi / a * np.array([[1.0, 1.0, 1.0]])
    v, A, S = orc.sk(x, k111)
    assert abs(abs(A[0]) - s.n_local) < 1e-8 * s.n_local
    assert abs(S[0] - s.n_local) < 1e-8 * s.n_local
    rng = np.random.default_rng(0)
    xr = rng.uniform(0, 20, size=(300, 3))
    k = rng.normal(size=(6, 3))
    _, _, S1 = orc.sk(xr, k)
    _, _, S2 = orc.sk(xr, -k)
    assert np.abs(S1 - S2).max() < 1e-10


def test_sk_gradient_fd(orc):
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 20, size=(200, 3))
    k = rng.normal(size=(10, 3)) * 2.0
    g = orc.sk_grad(x, k)
    h = 1e-6
    for j in (0, 17, 150):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[j, a] += h
            xm[j, a] -= h
            fd = (orc.sk(xp, k)[0] - orc.sk(xm, k)[0]) / (2 * h)
            assert abs(fd - g[j, a]) < 1e-6 * max(1.0, np.abs(g).max())


# ------------------------------------------------------------------ MLP
def _mlp_params(seed, H=64):
    rng = np.random.default_rng(seed)
    return [rng.normal(size=(H, 3)) * 0.5, rng.normal(size=H) * 0.1, rng.normal(size=(H, H)) / 8,
            rng.normal(size=H) * 0.1, rng.normal(size=H) / 8, 0.03]


def test_mlp_matches_torch_autograd(orc):
    torch = pytest.importorskip("torch")
    p = _mlp_params(2)
    rng = np.random.default_rng(3)
    s = rng.uniform(-1, 1, size=(16, 3))
    y = rng.uniform(-1, 1, size=16)
    V, dV = orc.mlp_forward(p, s)
    loss, grads = orc.mlp_loss_grad(p, s, y)
    tp = [torch.tensor(np.asarray(q, dtype=np.float64), requires_grad=True) for q in p]
    ts = torch.tensor(s, requires_grad=True)
    h1 = torch.tanh(ts @ tp[0].T + tp[1])
    h2 = torch.tanh(h1 @ tp[2].T + tp[3])
    out = h2 @ tp[4] + tp[5]
    assert np.abs(out.detach().numpy() - V).max() < 1e-12
    (gs,) = torch.autograd.grad(out.sum(), ts, retain_graph=True)
    assert np.abs(gs.numpy() - dV).max() < 1e-12
    L = torch.mean((out - torch.tensor(y)) ** 2)
    tg = torch.autograd.grad(L, tp)
    assert abs(L.item() - loss) < 1e-12
    for a, b in zip(tg, grads):
        assert np.abs(a.numpy() - np.asarray(b)).max() < 1e-12


# ------------------------------------------------------------------ FES reconstruction
def test_fes_1d_agrees_with_the_windowed_grid_on_the_reference_hills(orc):
    """tools/plot_hills.py's 1-D sum over the reference's own HILLS fixture agrees with the
    (independently pinned) uniform-grid deposit up to the grid's window truncation."""
    import os
    rows = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "HILLS_reference_test_cvset"), comments="#")
    x = np.linspace(0.0005, 0.9995, 1000)
    F = orc.fes_sum_1d(rows, x)
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "grid_reference_hills.npy"))
    assert np.abs(F - gold).max() <= 238 * 0.005 * np.exp(-8.0) + 1e-15


def test_fes_2d_window_truncation_bound(orc):
    """The 2-D windows drop only tails beyond ~4 sigma - d: bounded by h exp(-(4s - d)^2 / 2s^2) per hill."""
    rng = np.random.default_rng(21)
    n, lo, hi, nb = 60, np.array([0.0, -1.0]), np.array([1.0, 1.0]), np.array([80, 90])
    d = (hi - lo) / nb
    rows = np.column_stack([np.arange(n), rng.uniform(0.0, 1.0, n), rng.uniform(-1, 1, n),
                            rng.uniform(0.001, 0.005, n), np.full(n, 0.04)])
    x = np.linspace(lo[0] + d[0] / 2, hi[0] - d[0] / 2, nb[0])
    y = np.linspace(lo[1] + d[1] / 2, hi[1] - d[1] / 2, nb[1])
    F = orc.fes_sum_2d(rows, x, y, lo, d)
    X, Y = np.meshgrid(x, y, indexing="ij")
    full = sum(h * np.exp(-((X - a) ** 2 + (Y - b) ** 2) / (2 * s * s)) for _, a, b, h, s in rows)
    bound = np.sum(rows[:, 3]) * np.exp(-((4 * 0.04 - d.max()) ** 2) / (2 * 0.04 ** 2))
    assert np.abs(F - full).max() <= bound
    assert F.max() > 10 * bound


def test_fes_marginal_closed_forms(orc):
    """Constant F has a zero marginal; a separable F = a_i + b_j has F(CV1) = a - min(a)."""
    kT = 2.494
    assert np.abs(orc.fes_marginal(np.full((7, 9), 3.0), kT, 1, 0.1)).max() == 0.0
    rng = np.random.default_rng(5)
    a, b = rng.uniform(0, 10, 30), rng.uniform(0, 10, 40)
    F = a[:, None] + b[None, :]
    assert np.abs(orc.fes_marginal(F, kT, 1, 0.05) - (a - a.min())).max() <= 1e-12
    assert np.abs(orc.fes_marginal(F, kT, 0, 0.05) - (b - b.min())).max() <= 1e-12


# ------------------------------------------------------------------ DISTANCE
def test_distance_minimum_image_and_gradient(orc):
    """Across a periodic face the folded distance is the short one; the gradient is the FD
    of the distance (a move that keeps the pair away from the half-box fold)."""
    L = np.array([20.0, 15.0, 30.0])
    x = np.array([[0.5, 1.0, 2.0], [19.0, 14.0, 28.5]])
    r, d = orc.distance_cv(x, L, 0, 1)
    assert np.allclose(d, [-1.5, -2.0, -3.5]) and abs(r - np.sqrt(1.5 ** 2 + 2 ** 2 + 3.5 ** 2)) < 1e-15
    rng = np.random.default_rng(9)
    y = rng.uniform(0, 1, size=(2, 3)) * L * 0.3
    r0, d0 = orc.distance_cv(y, L, 0, 1)
    h = 1e-6
    for a, sgn in ((0, -1.0), (1, 1.0)):
        for k in range(3):
            yp, ym = y.copy(), y.copy()
            yp[a, k] += h
            ym[a, k] -= h
            fd = (orc.distance_cv(yp, L, 0, 1)[0] - orc.distance_cv(ym, L, 0, 1)[0]) / (2 * h)
            assert abs(fd - sgn * d0[k] / r0) < 1e-8


# ------------------------------------------------------------------ STRU_FACTOR (Debye)
def test_stru_factor_dimer_closed_form(orc):
    """Two atoms at r: S = 1 + sin(q r)/(q r) s(r^2) for each, with the cubic switch above 0.8 rc^2."""
    q, rc = 2.0, 5.0
    for r in (3.0, 4.8):
        x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        off, idx = np.array([0, 1, 2]), np.array([1, 0])
        s = 1.0 if r * r <= 0.8 * rc * rc else 1.0 - ((r * r - 0.8 * rc * rc) / (0.2 * rc * rc)) ** 3
        S = orc.sf_peratom(x, off, idx, 2, rc, q)
        assert np.allclose(S, 1.0 + np.sin(q * r) / (q * r) * s, rtol=0, atol=1e-15)


def test_stru_factor_gradient_fd(orc):
    """The exact gradient equals the FD of the CV, at group and non-group atoms (a small
    cluster, full brute-force lists, no periodic images)."""
    rng = np.random.default_rng(17)
    n = 14
    x = rng.uniform(0, 6.0, size=(n, 3))
    off = np.arange(n + 1) * (n - 1)
    idx = np.array([j for i in range(n) for j in range(n) if j != i])
    grp = (np.arange(n) % 3 != 0)
    q, rc = 1.7, 5.0
    G = orc.sf_grad(x, off, idx, n, np.zeros(0, int), rc, q, grp)
    h = 1e-6
    for k in (0, 1, 5, 9):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k, a] += h
            xm[k, a] -= h
            fd = (orc.sf_cv(xp, off, idx, n, rc, q, grp) - orc.sf_cv(xm, off, idx, n, rc, q, grp)) / (2 * h)
            assert abs(fd - G[k, a]) <= 1e-7 * max(1.0, np.abs(G).max())


def test_chem_weights_gaussian_and_map(orc):
    """Chem weights: exp(-(t - c)^2 / (2 sigma^2)) for unmapped types (1 at t = c, e^-1/2 one
    sigma away), the Chem_map value for mapped ones."""
    w = orc.chem_weights([0, 1, 2, 3], c_target=2.0, sigma=0.5, cmap={3: 0.25})
    assert w[2] == 1.0 and w[3] == 0.25
    assert abs(w[1] - np.exp(-2.0)) <= 1e-15 and abs(w[0] - np.exp(-8.0)) <= 1e-15
    assert orc.chem_weights([1])[0] == np.exp(-0.5)


@pytest.mark.parametrize("pair", [0, 1])
def test_chem_dimer_closed_form_and_scaling(orc, pair):
    """Two atoms with weights w_0, w_1 at r: S_i = 1 + w_0 w_1 g(r), g = sin(q r)/(q r) s(r^2)
    (STRU_FACTOR Chemical) or s(r^2) (WEIGHT_CHEM); scaling every weight by c scales S - 1 by
    c^2 in a cluster."""
    q, rc = 2.0, 5.0
    for r in (3.0, 4.7):
        x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        off, idx = np.array([0, 1, 2]), np.array([1, 0])
        s = 1.0 if r * r <= 0.8 * rc * rc else 1.0 - ((r * r - 0.8 * rc * rc) / (0.2 * rc * rc)) ** 3
        g = s if pair else np.sin(q * r) / (q * r) * s
        S = orc.sf_peratom(x, off, idx, 2, rc, q, aw=[0.3, 0.7], pair=pair)
        assert np.allclose(S, 1.0 + 0.21 * g, rtol=0, atol=1e-15)
    x, off, idx = _cluster(10, 23)
    aw = np.random.default_rng(3).uniform(0.2, 1.0, 10)
    S1 = orc.sf_peratom(x, off, idx, 10, rc, q, aw=aw, pair=pair)
    S3 = orc.sf_peratom(x, off, idx, 10, rc, q, aw=3.0 * aw, pair=pair)
    assert np.allclose(S3 - 1.0, 9.0 * (S1 - 1.0), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("pair", [0, 1])
def test_chem_gradients_fd(orc, pair):
    """AVE and COUNT gradients with unequal weights equal the central FD of the CV, at group
    and non-group atoms."""
    x, off, idx = _cluster(12, 29, L=5.5)
    n = 12
    aw = orc.chem_weights(np.arange(n) % 3 + 1, c_target=2.0, sigma=0.8, cmap={3: 0.4})
    grp = (np.arange(n) % 4 != 1)
    q, rc = 1.7, 5.0
    sw = dict(type="FERMI", r0=1.3, alpha=5.0)
    none = np.zeros(0, int)
    Ga = orc.sf_grad(x, off, idx, n, none, rc, q, grp, aw=aw, pair=pair)
    Gc = orc.sf_count_grad(x, off, idx, n, none, rc, q, sw, grp, aw=aw, pair=pair)
    h = 1e-6
    for k in (0, 1, 6, 11):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k, a] += h
            xm[k, a] -= h
            fa = (orc.sf_cv(xp, off, idx, n, rc, q, grp, aw, pair) - orc.sf_cv(xm, off, idx, n, rc, q, grp, aw, pair)) / (2 * h)
            fc = (orc.sf_count(xp, off, idx, n, rc, q, sw, grp, aw, pair)
                  - orc.sf_count(xm, off, idx, n, rc, q, sw, grp, aw, pair)) / (2 * h)
            assert abs(fa - Ga[k, a]) <= 1e-7 * max(1.0, np.abs(Ga).max())
            assert abs(fc - Gc[k, a]) <= 1e-7 * max(1.0, np.abs(Gc).max())


def test_switching_functions_derivatives_and_values(orc):
    """Each df is the FD of its f (every RATIONAL branch, the n = 2m one included, which is
    1 + x^m in the reference); FERMI and TANH are 1/2 at r_0, RATIONAL 1 at d_0 (m = 2n)."""
    cases = [dict(type="FERMI", r0=1.2, alpha=7.0), dict(type="TANH", r0=0.8, alpha=5.0),
             dict(type="RATIONAL", r0=1.25, d0=0.1, n=6, m=12), dict(type="RATIONAL", r0=1.1, d0=0.0, n=4, m=8),
             dict(type="RATIONAL", r0=1.3, d0=0.0, n=3, m=6), dict(type="RATIONAL", r0=0.9, d0=0.0, n=8, m=4),
             dict(type="RATIONAL", r0=1.4, d0=0.2, n=5, m=9)]
    for p in cases:
        for v in (0.3, 0.75, 1.6, 2.2):
            h = 1e-6
            fd = (orc.sw_f(p, v + h) - orc.sw_f(p, v - h)) / (2 * h)
            if p["type"] == "RATIONAL" and abs((v - p["d0"]) / p["r0"] - 1.0) < 1e-9 and 2 * p["n"] != p["m"]:
                # x = 1 in the general branch (0/0): the limit n (n - m) / (2 m r_0), which the
                # FD reaches only to its cancellation error
                n_, m_ = p["n"], p["m"]
                assert orc.sw_df(p, v) == 0.5 * (n_ * (n_ - m_)) / m_ / p["r0"]
                assert abs(fd - orc.sw_df(p, v)) <= 1e-4 * max(1.0, abs(fd)), (p, v)
                continue
            assert abs(fd - orc.sw_df(p, v)) <= 1e-6 * max(1.0, abs(fd)), (p, v)
    assert orc.sw_f(cases[0], 1.2) == 0.5 and orc.sw_f(cases[1], 0.8) == 0.5
    assert orc.sw_f(cases[2], 0.1) == 1.0
    assert orc.sw_f(cases[5], 0.9) == 2.0          # n = 2m: 1 + x^m at x = 1
    assert orc.sw_f(dict(type="STEP", r0=1.0), 0.99) == 1.0 and orc.sw_f(dict(type="STEP", r0=1.0), 1.0) == 0.0


def test_stru_factor_count_gradient_fd(orc):
    rng = np.random.default_rng(18)
    n = 12
    x = rng.uniform(0, 5.5, size=(n, 3))
    off = np.arange(n + 1) * (n - 1)
    idx = np.array([j for i in range(n) for j in range(n) if j != i])
    grp = (np.arange(n) % 4 != 0)
    sw = dict(type="FERMI", r0=1.5, alpha=4.0)
    q, rc = 1.7, 5.0
    G = orc.sf_count_grad(x, off, idx, n, np.zeros(0, int), rc, q, sw, grp)
    h = 1e-6
    for k in (0, 2, 7):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k, a] += h
            xm[k, a] -= h
            fd = (orc.sf_count(xp, off, idx, n, rc, q, sw, grp) - orc.sf_count(xm, off, idx, n, rc, q, sw, grp)) / (2 * h)
            assert abs(fd - G[k, a]) <= 1e-7 * max(1.0, np.abs(G).max())


# ---------------------------------------------------------------- STEINH L
def _cluster(n, seed, L=6.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, L, size=(n, 3))
    off = np.arange(n + 1) * (n - 1)
    idx = np.array([j for i in range(n) for j in range(n) if j != i])
    return x, off, idx


@pytest.mark.parametrize("l,q_ref", [(6, 0.574524), (4, 0.190941)])
def test_lq_perfect_fcc_known_answer(mdt, orc, l, q_ref):
    """Perfect FCC with STEP r_0 = 3.9: every atom weighs its 12 nearest by 1, all q_lm are
    equal, so LQ_l = q_l of the 12-neighbour shell (0.574524 for l = 6, 0.190941 for l = 4).
    With the default LINE f: MEAN_SOLID = LQ_l, NSOLID = N, FRAC = 1."""
    s = fcc_system(mdt, 4, sigma=0.0)
    x, off, idx, own = s.arrays()
    n = s.n_local
    step, line = dict(type="STEP", r0=3.9), dict(type="LINE")
    ql = orc.lq_peratom(x, off, idx, n, own, l, step)
    assert np.abs(ql - q_ref).max() <= 1e-6
    assert np.ptp(ql) <= 1e-12
    assert abs(orc.lq_cv(x, off, idx, n, own, l, step, line, mode="MEAN_SOLID") - ql.mean()) <= 1e-14
    assert orc.lq_cv(x, off, idx, n, own, l, step, line, mode="NSOLID") == n
    assert orc.lq_cv(x, off, idx, n, own, l, step, line, mode="FRAC") == 1.0
    fermi = dict(type="FERMI", r0=0.4, alpha=10.0)
    assert abs(orc.lq_cv(x, off, idx, n, own, l, step, fermi, mode="NSOLID")
               - n * orc.sw_f(fermi, ql[0])) <= 1e-9 * n


@pytest.mark.parametrize("mode", ["MEAN_SOLID", "NSOLID", "FRAC"])
@pytest.mark.parametrize("l", [4, 6])
def test_lq_gradient_fd_cluster(orc, mode, l):
    """The exact gradient equals the central FD of the CV at group and non-group atoms
    (a 13-atom cluster, all pairs listed, RATIONAL weights in r, FERMI f on LQ_l)."""
    x, off, idx = _cluster(13, 40 + l)
    n = 13
    grp = np.arange(n) % 4 != 1
    swr = dict(type="RATIONAL", r0=2.6, d0=0.0, n=6, m=12)
    swcv = dict(type="FERMI", r0=0.35, alpha=8.0)
    own = np.zeros(0, int)
    G = orc.lq_grad(x, off, idx, n, own, l, swr, swcv, 1e-6, mode, grp)
    h = 1e-6
    scale = max(1e-3, np.abs(G).max())
    for k in (0, 1, 5, 10):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k, a] += h
            xm[k, a] -= h
            fd = (orc.lq_cv(xp, off, idx, n, own, l, swr, swcv, 1e-6, mode, grp)
                  - orc.lq_cv(xm, off, idx, n, own, l, swr, swcv, 1e-6, mode, grp)) / (2 * h)
            assert abs(fd - G[k, a]) <= 2e-7 * scale, (mode, k, a, fd, G[k, a])
    # translation invariance (no periodic images: the forces sum to zero)
    assert np.abs(G.sum(axis=0)).max() <= 1e-12 * max(1.0, np.abs(G).max() * n)


def test_lq_gradient_fd_periodic_with_ghosts(mdt, orc):
    """On a periodic system the ghost entries fold onto their owners: FD with ghost-consistent
    moves (Mover) equals the gradient, and the forces sum to zero."""
    from helpers import Mover
    s = fcc_system(mdt, 3, sigma=0.25, seed=44)
    x, off, idx, own = s.arrays()
    n = s.n_local
    grp = (x[:n, 2] < x[:n, 2].mean())
    swr = dict(type="RATIONAL", r0=3.0, d0=0.0, n=6, m=12)
    swcv = dict(type="TANH", r0=0.3, alpha=6.0)
    G = orc.lq_grad(x, off, idx, n, own, 6, swr, swcv, 1e-6, "MEAN_SOLID", grp)
    mv = Mover(x, n, own)
    h = 1e-6
    for k in (0, 17, 60):
        for a in range(3):
            xl = x[:n].copy()
            xl[k, a] += h
            cp = orc.lq_cv(mv.set_local(xl), off, idx, n, own, 6, swr, swcv, 1e-6, "MEAN_SOLID", grp)
            xl[k, a] -= 2 * h
            cm = orc.lq_cv(mv.set_local(xl), off, idx, n, own, 6, swr, swcv, 1e-6, "MEAN_SOLID", grp)
            assert abs((cp - cm) / (2 * h) - G[k, a]) <= 2e-7 * np.abs(G).max(), (k, a)
    assert np.abs(G.sum(axis=0)).max() <= 1e-12 * np.abs(G).max() * n


@pytest.mark.parametrize("mode", ["MEAN_SOLID", "NSOLID", "FRAC"])
def test_lq_c_restatement_matches_python(orc, mode):
    """orc_lq (C, OpenMP) equals the Python STEINH L oracle, value and gradient, on a
    cluster with a group and a RATIONAL r weight, and on the periodic lattice deck."""
    x, off, idx = _cluster(16, 41, L=6.5)
    n = 16
    grp = (np.arange(n) % 5 != 2)
    swr = dict(type="RATIONAL", r0=3.0, d0=0.0, n=6, m=12)
    swc = dict(type="FERMI", r0=0.3, alpha=12.0)
    none = np.zeros(0, int)
    cv_p = orc.lq_cv(x, off, idx, n, none, 6, swr, swc, 1e-6, mode, grp)
    G_p = orc.lq_grad(x, off, idx, n, none, 6, swr, swc, 1e-6, mode, grp)
    cv_c, G_c = orc.lq_c(x, off, idx, n, none, 6, swr, swc, 1e-6, mode, grp)
    assert abs(cv_c - cv_p) <= 1e-12 * max(1.0, abs(cv_p))
    assert np.abs(G_c - G_p).max() <= 1e-11 * max(1.0, np.abs(G_p).max())


def test_lq_c_periodic_matches_python(mdt, orc):
    s = fcc_system(mdt, 3, sigma=0.2, seed=77)
    x, off, idx, own = s.arrays()
    n = s.n_local
    swr = dict(type="RATIONAL", r0=3.0, d0=0.0, n=6, m=12)
    swc = dict(type="TANH", r0=0.35, alpha=9.0)
    cv_p = orc.lq_cv(x, off, idx, n, own, 4, swr, swc, 1e-6, "MEAN_SOLID")
    G_p = orc.lq_grad(x, off, idx, n, own, 4, swr, swc, 1e-6, "MEAN_SOLID")
    cv_c, G_c = orc.lq_c(x, off, idx, n, own, 4, swr, swc, 1e-6, "MEAN_SOLID")
    assert abs(cv_c - cv_p) <= 1e-12 * max(1.0, abs(cv_p))
    assert np.abs(G_c - G_p).max() <= 1e-11 * max(1.0, np.abs(G_p).max())
