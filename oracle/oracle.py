"""ctypes front end of the FP64 CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, never by the shipped library.  Every function
here calls the C restatement (numpy only marshals arrays), except the parts at the
end: the FES reconstruction (numpy, as the reference's tools/plot_hills.py is), and
DISTANCE, STRU_FACTOR, the switching functions and STEINH L, restated as plain Python
loops over atoms and list entries for the small systems the tests use (Y_lm and its
gradient still come from the C restatement).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
_L = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-B", "-C", HERE], check=True, capture_output=True)


def lib():
    global _L
    if _L is None:
        build()
        _L = C.CDLL(LIB)
        P = C.c_void_p
        I, D = C.c_int, C.c_double
        for name, res, args in [
            ("orc_num_threads", I, []),
            ("orc_ylm", None, [I, D, D, D, P]),
            ("orc_ylm_poly", None, [I, D, D, D, P]),
            ("orc_ylm_grad", None, [I, D, D, D, P]),
            ("orc_ylm_coef", None, [P]),
            ("orc_select", None, [I, P, P, P, P, D, I, D, P, P, P, P, P]),
            ("orc_steinhardt", None, [I, I, I, P, P, P, P]),
            ("orc_kahan_mean", D, [I, P, P, D]),
            ("orc_ld_local", None, [I, I, I, I, P, P, P, P, P, P]),
            ("orc_grad_from_coeff", None, [I, I, I, I, P, P, P, P, P, P, P]),
            ("orc_q_coeff", None, [I, I, P, P, D, D, P]),
            ("orc_ld_coeff", None, [I, I, I, I, P, P, P, P, P, D, D, P]),
            ("orc_sk", D, [I, P, I, P, P, P]),
            ("orc_sk_grad", None, [I, P, I, P, P, D, P]),
            ("orc_grid_radius", I, [D, D, D, I]),
            ("orc_grid_deposit", None, [I, P, P, P, P, P, D, D]),
            ("orc_grid_eval", D, [I, P, P, P, P, P, P]),
            ("orc_wt_height", D, [I, P, P, P, P, P, D, D, D, D, I]),
            ("orc_gridt1_deposit", None, [I, P, P, P, P, P, D, D]),
            ("orc_gridt1_eval", D, [I, P, P, P, P, P, P]),
            ("orc_hills_sum", None, [I, I, P, P, P, I, P, P, P]),
            ("orc_mlp_forward", None, [I, I, P, P, P, P, P, D, I, P, P, P]),
            ("orc_mlp_loss_grad", D, [I, I, P, P, P, P, P, D, I, P, P, P, P, P, P, P, P]),
            ("orc_lq", D, [I, I, P, P, P, P, P, P, P, D, I, P]),
        ]:
            f = getattr(_L, name)
            f.restype = res
            f.argtypes = args
    return _L


def threads():
    """Threads the oracle's OpenMP loops run on."""
    return lib().orc_num_threads()


def host_cpu():
    """Model name of the host CPU and its logical core count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def _p(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- Y_lm
def ylm(l, d):
    out = np.zeros(2 * (l + 1))
    lib().orc_ylm(l, float(d[0]), float(d[1]), float(d[2]), _p(out))
    return out[0::2] + 1j * out[1::2]


def ylm_poly(l, d):
    out = np.zeros(2 * (l + 1))
    lib().orc_ylm_poly(l, float(d[0]), float(d[1]), float(d[2]), _p(out))
    return out[0::2] + 1j * out[1::2]


def ylm_grad(l, d):
    """(l+1, 3) complex: d Y_lm / d r_a."""
    out = np.zeros(6 * (l + 1))
    lib().orc_ylm_grad(l, float(d[0]), float(d[1]), float(d[2]), _p(out))
    o = out.reshape(l + 1, 3, 2)
    return o[..., 0] + 1j * o[..., 1]


# ---------------------------------------------------------------- selection / Steinhardt
class Env:
    """Selection of every local atom (centers = all locals, or a group)."""

    def __init__(self, x_all, off, idx, owner, n_local, rc, N, group=None):
        self.x = _f64(x_all).reshape(-1, 3)
        self.off, self.idx, self.owner = _i32(off), _i32(idx), _i32(owner)
        self.n_local, self.rc, self.N = n_local, rc, N
        if group is None:
            centers = np.arange(n_local, dtype=np.int32)
        else:
            centers = np.nonzero(np.asarray(group))[0].astype(np.int32)
        self.centers = centers
        nc = centers.size
        # per-center CSR (the oracle takes lists in center order)
        if group is None:
            offs, cidx = self.off, self.idx
        else:
            offs = np.zeros(nc + 1, np.int32)
            lst = []
            for k, i in enumerate(centers):
                seg = self.idx[self.off[i]:self.off[i + 1]]
                lst.append(seg)
                offs[k + 1] = offs[k] + seg.size
            cidx = np.concatenate(lst).astype(np.int32) if lst else np.zeros(0, np.int32)
        self.sel_idx = np.zeros((nc, N), np.int32)
        self.sel_pos = np.zeros((nc, N), np.int32)
        self.sel_d = np.zeros((nc, N, 4))
        self.cnt = np.zeros(nc, np.int32)
        self.nsel = np.zeros(nc, np.int32)
        lo = self.x.min(axis=0)
        L = np.ptp(self.x[:n_local], axis=0)
        box_sum2 = float(np.sum(L) ** 2)
        lib().orc_select(nc, _p(centers), _p(offs), _p(cidx), _p(self.x), rc * rc, N, box_sum2,
                         _p(self.sel_idx), _p(self.sel_pos), _p(self.sel_d), _p(self.cnt), _p(self.nsel))
        self._full = group is None
        self._st = {}

    def owner_of(self, j):
        j = np.asarray(j)
        return np.where(j < self.n_local, j, self.owner[np.clip(j - self.n_local, 0, None)])

    def steinhardt(self, l):
        """(q_lm, q_l) of every center; computed once per degree (the selection is fixed)."""
        if l not in self._st:
            nc = self.centers.size
            q = np.zeros(nc * 2 * (l + 1))
            ql = np.zeros(nc)
            lib().orc_steinhardt(l, nc, self.N, _p(self.sel_d), _p(self.nsel), _p(q), _p(ql))
            self._st[l] = (q.reshape(nc, 2 * (l + 1)), ql)
        q, ql = self._st[l]
        return q.copy(), ql.copy()

    def q_ave(self, l):
        _, ql = self.steinhardt(l)
        return lib().orc_kahan_mean(ql.size, _p(ql), None, float(ql.size))

    def ld(self, l):
        assert self._full, "local Q oracle runs with centers == all local atoms"
        q, _ = self.steinhardt(l)
        qb = np.zeros_like(q)
        lq = np.zeros(self.n_local)
        lib().orc_ld_local(l, self.n_local, self.N, self.n_local, _p(self.owner), _p(self.sel_idx),
                           _p(self.nsel), _p(q), _p(qb), _p(lq))
        return qb, lq

    def dQ_dx(self, l, w=1.0):
        """w * d Q_l.AVE / dx for local atoms (n_local, 3); Q_l.AVE averages over the
        centers (all local atoms, or the group), and the gradient covers every atom."""
        q, ql = self.steinhardt(l)
        nc = self.centers.size
        K = np.zeros_like(q)
        lib().orc_q_coeff(l, nc, _p(q), _p(ql), float(nc), w, _p(K))
        out = np.zeros((self.n_local, 3))
        lib().orc_grad_from_coeff(l, nc, self.N, self.n_local, _p(self.owner), _p(self.sel_idx),
                                  _p(self.sel_d), _p(self.nsel), _p(K), _p(out),
                                  None if self._full else _p(self.centers))
        return out

    def dLQ_dx(self, l, w=1.0):
        qb, lq = self.ld(l)
        K = np.zeros_like(qb)
        lib().orc_ld_coeff(l, self.n_local, self.N, self.n_local, _p(self.owner), _p(self.sel_idx),
                           _p(self.nsel), _p(qb), _p(lq), float(self.n_local), w, _p(K))
        out = np.zeros((self.n_local, 3))
        lib().orc_grad_from_coeff(l, self.n_local, self.N, self.n_local, _p(self.owner), _p(self.sel_idx),
                                  _p(self.sel_d), _p(self.nsel), _p(K), _p(out), None)
        return out


def kahan_mean(v):
    v = _f64(v)
    return lib().orc_kahan_mean(v.size, _p(v), None, float(v.size))


# ---------------------------------------------------------------- structure factor
def sk(x, k):
    x, k = _f64(x).reshape(-1, 3), _f64(k).reshape(-1, 3)
    A = np.zeros(2 * k.shape[0])
    S = np.zeros(k.shape[0])
    v = lib().orc_sk(x.shape[0], _p(x), k.shape[0], _p(k), _p(A), _p(S))
    return v, A[0::2] + 1j * A[1::2], S


def sk_grad(x, k, w=1.0, A=None):
    """w dCV/dr_j (n, 3); A = the amplitudes from sk() when already computed."""
    x, k = _f64(x).reshape(-1, 3), _f64(k).reshape(-1, 3)
    if A is None:
        _, A, _ = sk(x, k)
    Ai = np.zeros(2 * k.shape[0])
    Ai[0::2], Ai[1::2] = A.real, A.imag
    out = np.zeros_like(x)
    lib().orc_sk_grad(x.shape[0], _p(x), k.shape[0], _p(k), _p(Ai), w, _p(out))
    return out


# ---------------------------------------------------------------- grid / hills
class Grid:
    def __init__(self, lo, hi, nb):
        self.lo, self.hi = _f64(lo), _f64(hi)
        self.nb = _i32(nb)
        self.D = self.lo.size
        self.g = np.zeros(int(np.prod(self.nb)))

    def deposit(self, s, w, sigma):
        s = _f64(s)
        lib().orc_grid_deposit(self.D, _p(self.lo), _p(self.hi), _p(self.nb), _p(self.g), _p(s), w, sigma)

    def eval(self, s):
        s = _f64(s)
        dV = np.zeros(self.D)
        V = lib().orc_grid_eval(self.D, _p(self.lo), _p(self.hi), _p(self.nb), _p(self.g), _p(s), _p(dV))
        return V, dV

    def wt_height(self, s, h0, biasf, T, kB, wt):
        s = _f64(s)
        return lib().orc_wt_height(self.D, _p(self.lo), _p(self.hi), _p(self.nb), _p(self.g), _p(s),
                                   h0, biasf, T, kB, int(wt))


class GridT1(Grid):
    """Sparse-hash bias (Gaussian_Hill_type 1) as a dense grid of D <= 3 dimensions,
    restated from GH_t1_sparseHash.cu (see orc_gridt1_* in oracle.cpp)."""

    def deposit(self, s, w, sigma):
        s = _f64(s)
        lib().orc_gridt1_deposit(self.D, _p(self.lo), _p(self.hi), _p(self.nb), _p(self.g), _p(s), w, sigma)

    def eval(self, s):
        s = _f64(s)
        dV = np.zeros(self.D)
        V = lib().orc_gridt1_eval(self.D, _p(self.lo), _p(self.hi), _p(self.nb), _p(self.g), _p(s), _p(dV))
        return V, dV

    def wt_height(self, s, h0, biasf, T, kB, wt):
        V = self.eval(s)[0] if wt else 0.0
        return h0 * np.exp(-V / (T * kB * (biasf - 1.0)))


def hills_sum(centers, heights, sigma, s, hill_sigma=None):
    """Exact sum of Gaussians at points s (np, D): V (np,) and dV/ds (np, D).  sigma: widths
    per dimension; hill_sigma (per hill, optional): hill h has widths sigma * hill_sigma[h] /
    sigma[0] (a hill replayed from a HILLS file with its own sigma)."""
    c = _f64(centers)
    D = c.shape[1]
    h, sg, s = _f64(heights), _f64(sigma), _f64(np.atleast_2d(s))
    V = np.zeros(s.shape[0])
    dV = np.zeros_like(s)
    if hill_sigma is None:
        lib().orc_hills_sum(D, h.size, _p(c), _p(h), _p(sg), s.shape[0], _p(s), _p(V), _p(dV))
        return V, dV
    hs = _f64(hill_sigma)
    for v in np.unique(hs):
        k = np.nonzero(hs == v)[0]
        Vk, dVk = np.zeros_like(V), np.zeros_like(dV)
        sk_ = _f64(sg * (v / sg[0]))
        ck, hk = _f64(c[k]), _f64(h[k])
        lib().orc_hills_sum(D, hk.size, _p(ck), _p(hk), _p(sk_), s.shape[0], _p(s), _p(Vk), _p(dVk))
        V += Vk
        dV += dVk
    return V, dV


# ---------------------------------------------------------------- MLP
def mlp_forward(params, s):
    W1, b1, W2, b2, W3, b3 = [_f64(p) for p in params]
    s = _f64(np.atleast_2d(s))
    H, Din = W1.shape
    V = np.zeros(s.shape[0])
    dV = np.zeros_like(s)
    lib().orc_mlp_forward(Din, H, _p(W1), _p(b1), _p(W2), _p(b2), _p(W3), float(np.asarray(b3).reshape(-1)[0]), s.shape[0], _p(s),
                          _p(V), _p(dV))
    return V, dV


def mlp_loss_grad(params, s, y):
    W1, b1, W2, b2, W3, b3 = [_f64(p) for p in params]
    s, y = _f64(np.atleast_2d(s)), _f64(y)
    H, Din = W1.shape
    g = [np.zeros_like(W1), np.zeros_like(b1), np.zeros_like(W2), np.zeros_like(b2), np.zeros_like(W3)]
    gb3 = np.zeros(1)
    loss = lib().orc_mlp_loss_grad(Din, H, _p(W1), _p(b1), _p(W2), _p(b2), _p(W3), float(np.asarray(b3).reshape(-1)[0]), s.shape[0],
                                   _p(s), _p(y), _p(g[0]), _p(g[1]), _p(g[2]), _p(g[3]), _p(g[4]), _p(gb3))
    return loss, g + [gb3[0]]


# ------------------------------------------------------------------ FES reconstruction
# Restated from the reference's post-processing script tools/plot_hills.py in its own
# language (numpy).  Hills rows: step, s..., height, sigma.
def fes_sum_1d(hills, x):
    """:158-172: every hill at every point, summed over hills in file order."""
    hills = np.atleast_2d(np.asarray(hills, np.float64))
    c, h, s = hills[:, 1, None], hills[:, 2, None], hills[:, 3, None]
    return np.sum(h * np.exp(-((np.asarray(x)[None, :] - c) ** 2) / (2 * s ** 2)), axis=0)


def fes_sum_2d(hills, x, y, lo, d):
    """:174-223: each hill over the bins of its window int((c -+ 4 sigma - lo)/d), clamped."""
    hills = np.atleast_2d(np.asarray(hills, np.float64))
    X, Y = np.meshgrid(x, y, indexing="ij")
    F = np.zeros_like(X)
    n0, n1 = X.shape
    for _, c1, c2, h, s in hills:
        cut, t2 = 4.0 * s, 2.0 * s ** 2
        r0, r1 = max(0, int((c1 - cut - lo[0]) / d[0])), min(n0 - 1, int((c1 + cut - lo[0]) / d[0]))
        q0, q1 = max(0, int((c2 - cut - lo[1]) / d[1])), min(n1 - 1, int((c2 + cut - lo[1]) / d[1]))
        if r0 <= r1 and q0 <= q1:
            Xl, Yl = X[r0:r1 + 1, q0:q1 + 1], Y[r0:r1 + 1, q0:q1 + 1]
            F[r0:r1 + 1, q0:q1 + 1] += h * np.exp(-((Xl - c1) ** 2 + (Yl - c2) ** 2) / t2)
    return F


def fes_marginal(F, kT, axis, bin_size):
    """:236-262: -kT ln(sum exp(-F/kT) * bin) along axis, minimum shifted to 0."""
    m = -kT * np.log(np.sum(np.exp(-np.asarray(F) / kT), axis=axis) * bin_size)
    return m - np.min(m)


# ------------------------------------------------------------------ DISTANCE
def distance_cv(x, L, a1, a2):
    """src/cvs/CVs_DISTANCE.cu:55-121: r = |x_a2 - x_a1| with each component folded by the
    half-box rule of delta_x (> L/2: -L; < -L/2: +L); returns r and the folded vector d, so
    dr/dx_a1 = -d/r and dr/dx_a2 = d/r (get_dcvdx)."""
    d = np.asarray(x[a2], np.float64) - np.asarray(x[a1], np.float64)
    for k in range(3):
        if d[k] > L[k] / 2:
            d[k] -= L[k]
        elif d[k] < -L[k] / 2:
            d[k] += L[k]
    return float(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])), d


# ------------------------------------------------------------------ STRU_FACTOR (Debye, AVE)
def _sf_switch(r2, rc2, r):
    """s and ds/dr of src/cvs/CVs_Strufactor.cu:1112-1114,1178-1184 (r_on = 0.8 rc^2)."""
    r_on = 0.8 * rc2
    if r2 > r_on:
        u = (r2 - r_on) / (rc2 - r_on)
        return 1.0 - u * u * u, -3 * (u * u) * (2.0 * r) / (rc2 - r_on)
    return 1.0, 0.0


def chem_weights(types, c_target=0.0, sigma=1.0, cmap=None):
    """Per-atom chemical weight of Stru_fact_chem_env (CVs_Strufactor_chem.cu:44-66): the
    Chem_map value of the atom's type, else exp(-(type - c_target)^2 / (2 sigma^2))."""
    cmap = cmap or {}
    return np.array([cmap[t] if t in cmap else np.exp(-((t - c_target) ** 2) / (2.0 * sigma * sigma))
                     for t in np.asarray(types, np.int64)], np.float64)


def _sf_g(r2, rc2, q, pair):
    """Pair term g and dg/dr: pair 0, sin(q r)/(q r) s(r^2) (STRU_FACTOR,
    structure_factor_cv_kernel :1082-1120); pair 1, s(r^2) (WEIGHT_CHEM,
    Weighted_chem_pair_cv_kernel, CV_Weighted_Chemical_Pair_Order.cu:647-694)."""
    r = np.sqrt(r2)
    s, ds = _sf_switch(r2, rc2, r)
    if pair:
        return s, ds
    th = q * r
    return np.sin(th) / th * s, (np.cos(th) / th - np.sin(th) / th ** 2) * s * q + ds * np.sin(th) / th


def _aw_of(aw, n_all):
    return np.ones(n_all) if aw is None else np.asarray(aw, np.float64)


def sf_peratom(x_all, off, idx, n, rc, q, aw=None, pair=0):
    """S_i = 1 + w_i sum_j w_j g(r_ij) over list entries with 1e-12 <= r^2 <= rc^2
    (get_environment_Strufactor :1014-1080; the chemical form multiplies the pair term by
    the two atoms' weights, cv_kernel_structure_factor_chem, CVs_Strufactor_chem.cu:151-191).
    aw: weights of every atom (locals then ghosts), None = 1."""
    x = np.asarray(x_all, np.float64)
    w = _aw_of(aw, x.shape[0])
    rc2 = rc * rc
    S = np.zeros(n)
    for i in range(n):
        acc = 0.0
        for e in range(off[i], off[i + 1]):
            j = idx[e]
            dx, dy, dz = x[j] - x[i]
            r2 = dx * dx + dy * dy + dz * dz
            if r2 > rc2 or r2 < 1e-12:
                continue
            g, _ = _sf_g(r2, rc2, q, pair)
            acc += w[j] * g
        S[i] = w[i] * acc + 1.0
    return S


def sf_cv(x_all, off, idx, n, rc, q, group=None, aw=None, pair=0):
    """compute_cv_AVE (:642-656): sum of S_i over the group / N_group."""
    S = sf_peratom(x_all, off, idx, n, rc, q, aw, pair)
    g = np.ones(n, bool) if group is None else np.asarray(group, bool)
    return float(np.sum(S[g]) / g.sum())


def sf_grad(x_all, off, idx, n, own, rc, q, group=None, aw=None, pair=0):
    """Exact dCV/dx of sf_cv for the n local atoms: pair (k, j) enters S_k for k in the group
    and S_j for j in it (a ghost j counts as its owner), with weight w_k w_j."""
    x = np.asarray(x_all, np.float64)
    w = _aw_of(aw, x.shape[0])
    g = np.ones(n, bool) if group is None else np.asarray(group, bool)
    rc2 = rc * rc
    out = np.zeros((n, 3))
    for k in range(n):
        for e in range(off[k], off[k + 1]):
            j = idx[e]
            d = x[j] - x[k]
            r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            if r2 > rc2 or r2 < 1e-12:
                continue
            r = np.sqrt(r2)
            _, fp = _sf_g(r2, rc2, q, pair)
            jo = j if j < n else own[j - n]
            wt = (float(g[k]) + float(g[jo])) * w[k] * w[j]
            out[k] -= wt * fp * d / r
    return out / g.sum()


# ------------------------------------------------------------------ switching functions
def sw_f(p, v):
    """include/utils/zqc_switch_function.h:216-265 (f); p = dict(type, r0, d0, alpha, n, m),
    type one of LINE STEP FERMI TANH RATIONAL."""
    t = p["type"]
    if t == "STEP":
        return 1.0 if v < p["r0"] else 0.0
    if t == "FERMI":
        return 1.0 / (1.0 + np.exp(-p["alpha"] * (v - p["r0"])))
    if t == "TANH":
        return 0.5 * (1.0 + np.tanh(p["alpha"] * (v - p["r0"])))
    if t == "RATIONAL":
        n, m = p["n"], p["m"]
        x = (v - p["d0"]) / p["r0"]
        if 2 * n == m:
            return 1.0 / (1.0 + x ** n)
        if n == 2 * m:
            return 1.0 + x ** m
        return (1 - x ** n) / (1 - x ** m)
    return 1.0


def sw_df(p, v):
    """:230-337 (df/dx): the reference's closed forms per type, except the x = 1 limit of the
    general RATIONAL branch, taken with the sign the derivative has."""
    t = p["type"]
    if t == "FERMI":
        e = np.exp(-p["alpha"] * (v - p["r0"]))
        return p["alpha"] * e / (1.0 + e) ** 2
    if t == "TANH":
        return 0.5 * p["alpha"] * (1.0 - np.tanh(p["alpha"] * (v - p["r0"])) ** 2)
    if t == "RATIONAL":
        n, m, r0 = p["n"], p["m"], p["r0"]
        x = (v - p["d0"]) / r0
        if 2 * n == m:
            return -n * x ** (n - 1) / (1.0 + x ** n) ** 2 / r0
        if n == 2 * m:
            return m * x ** (m - 1) / r0
        xm = x ** m
        if abs(1.0 - xm) < 1e-12:
            # the limit of the derivative at x = 1 (the reference's :325-327 has the sign flipped)
            return 0.5 * (n * (n - m)) / m / r0
        return ((n - m) * x ** (n - 1) * xm + m * x ** (m - 1) - n * x ** (n - 1)) / (1 - xm) ** 2 / r0
    return 0.0


def sf_count(x_all, off, idx, n, rc, q, sw, group=None, aw=None, pair=0):
    """compute_cv_COUNT (CVs_Strufactor.cu:658-674): sum over the group of sw(S_i)."""
    S = sf_peratom(x_all, off, idx, n, rc, q, aw, pair)
    g = np.ones(n, bool) if group is None else np.asarray(group, bool)
    return float(sum(sw_f(sw, S[i]) for i in range(n) if g[i]))


def sf_count_grad(x_all, off, idx, n, own, rc, q, sw, group=None, aw=None, pair=0):
    """Exact dCOUNT/dx: pair (k, j) carries ([k in G] sw'(S_k) + [j in G] sw'(S_j)) w_k w_j."""
    x = np.asarray(x_all, np.float64)
    w = _aw_of(aw, x.shape[0])
    S = sf_peratom(x_all, off, idx, n, rc, q, aw, pair)
    g = np.ones(n, bool) if group is None else np.asarray(group, bool)
    rc2 = rc * rc
    out = np.zeros((n, 3))
    for k in range(n):
        for e in range(off[k], off[k + 1]):
            j = idx[e]
            d = x[j] - x[k]
            r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            if r2 > rc2 or r2 < 1e-12:
                continue
            r = np.sqrt(r2)
            _, fp = _sf_g(r2, rc2, q, pair)
            jo = j if j < n else own[j - n]
            wt = ((sw_df(sw, S[k]) if g[k] else 0.0) + (sw_df(sw, S[jo]) if g[jo] else 0.0)) * w[k] * w[j]
            out[k] -= wt * fp * d / r
    return out


# ------------------------------------------------------------------ STEINH L (switching-weighted local Q)
LQ_MODES = ("MEAN_SOLID", "NSOLID", "FRAC")


def _lq_state(x_all, off, idx, n, own, l, swr, eps, group):
    """Per atom a over its list entries with sigma(r) >= eps (get_environment_Steinhardt_LocalQ,
    src/cvs/CVs_Steinhardt_Local.cu:1210-1285): S_a = sum sigma, q_lm(a) = sum sigma Y_lm / S_a
    (steinhardt_Local_cv_get_qlm :1310-1452; zero when no entry counts or S_a < 1e-12).  For a
    group atom i: LQ_lm(i) = (q_lm(i) + sum sigma_ij q_lm(j)) / (S_i + 1) with q_lm of a ghost j
    its owner's, and LQ_l(i) = sqrt(4 pi/(2l+1) sum_m |LQ_lm|^2) (steinhardt_Local_cv_get_LQl
    :1490-1579)."""
    x = np.asarray(x_all, np.float64)
    g = np.ones(n, bool) if group is None else np.asarray(group, bool)
    S = np.zeros(n)
    q = np.zeros((n, l + 1), complex)
    live = np.zeros(n, bool)
    nb = [[] for _ in range(n)]
    for a in range(n):
        acc = np.zeros(l + 1, complex)
        for e in range(off[a], off[a + 1]):
            j = idx[e]
            d = x[j] - x[a]
            r2 = float(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
            if r2 < 1e-12:
                continue
            r = float(np.sqrt(r2))
            s = sw_f(swr, r)
            if s < eps:
                continue
            o = j if j < n else int(own[j - n])
            nb[a].append((o, s, sw_df(swr, r), d, r))
            S[a] += s
            acc += s * ylm(l, d)
        live[a] = len(nb[a]) > 0 and S[a] >= 1e-12
        q[a] = acc * (1.0 / S[a]) if live[a] else 0.0
    LQ = np.zeros((n, l + 1), complex)
    ql = np.zeros(n)
    cm = np.where(np.arange(l + 1) > 0, 2.0, 1.0)
    for i in range(n):
        if not (g[i] and live[i]):
            continue
        acc = q[i].copy()
        for (o, s, _, _, _) in nb[i]:
            acc += s * q[o]
        LQ[i] = acc * (1.0 / (S[i] + 1.0))
        ql[i] = float(np.sqrt(np.sum(cm * np.abs(LQ[i]) ** 2) * 4.0 * np.pi / (2 * l + 1)))
    return dict(g=g, S=S, q=q, live=live, nb=nb, LQ=LQ, ql=ql, cm=cm)


def _lq_reduce(st, swcv, mode):
    """reduce_lq_sw_sums / compute_cv_{AVE,NSOLID,FRAC} (:694-745): over local group atoms,
    Sq = sum q f(q), Sf = sum f(q); MEAN_SOLID = Sq/Sf (0 if Sf = 0), NSOLID = Sf,
    FRAC = Sf / N_group."""
    g, ql = st["g"], st["ql"]
    f = np.array([sw_f(swcv, v) for v in ql[g]])
    Sq, Sf, N = float(np.sum(ql[g] * f)), float(np.sum(f)), int(g.sum())
    if mode == "MEAN_SOLID":
        cv = Sq / Sf if Sf != 0.0 else 0.0
    elif mode == "NSOLID":
        cv = Sf
    else:
        cv = Sf / N if N > 0 else 0.0
    return cv, Sf, N


_SW_CODE = {"LINE": 0, "STEP": 1, "FERMI": 2, "TANH": 3, "RATIONAL": 4}


def _sw_vec(p):
    return np.array([_SW_CODE[p["type"]], p.get("r0", 0.0), p.get("d0", 0.0), p.get("alpha", 0.0),
                     p.get("n", 0), p.get("m", 0)], np.float64)


def lq_c(x_all, off, idx, n, own, l, swr, swcv, eps=1e-6, mode="MEAN_SOLID", group=None, grad=True):
    """(CV, dCV/dx or None): lq_cv and lq_grad restated in C with OpenMP over atoms
    (orc_lq), for systems the Python loops take too long on."""
    x = _f64(x_all).reshape(-1, 3)
    o, ix, ow = _i32(off), _i32(idx), _i32(own if len(own) else np.zeros(1, np.int32))
    g = None if group is None else np.ascontiguousarray(group, np.uint8)
    G = np.zeros((n, 3)) if grad else None
    pr, pc = _sw_vec(swr), _sw_vec(swcv)          # held until the call returns
    cv = lib().orc_lq(l, n, _p(x), _p(o), _p(ix), _p(ow), _p(g), _p(pr), _p(pc), eps, LQ_MODES.index(mode), _p(G))
    return cv, G


def lq_cv(x_all, off, idx, n, own, l, swr, swcv, eps=1e-6, mode="MEAN_SOLID", group=None):
    """The STEINH L CV value; swr, swcv as for sw_f (dict(type="LINE") for the default)."""
    st = _lq_state(x_all, off, idx, n, own, l, swr, eps, group)
    return _lq_reduce(st, swcv, mode)[0]


def lq_peratom(x_all, off, idx, n, own, l, swr, eps=1e-6, group=None):
    """Per-atom LQ_l (0 outside the group and where no entry counts)."""
    return _lq_state(x_all, off, idx, n, own, l, swr, eps, group)["ql"]


def lq_grad(x_all, off, idx, n, own, l, swr, swcv, eps=1e-6, mode="MEAN_SOLID", group=None):
    """Exact dCV/dx of lq_cv for the n local atoms (a ghost entry counts for its owner).
    Adjoint chain: G_i = dCV/dLQ_lm(i) = w_i 4pi/(2l+1) c_m LQ_lm / LQ_l with w_i = dCV/dLQ_l(i)
    (MEAN_SOLID (f + f'(q - CV)) / Sf, NSOLID f', FRAC f' / N); K_i = G_i / (S_i + 1);
    H_a = dCV/dq_lm(a) = K_a + sum over centres i that count a of sigma_ia K_i (scattered
    here, exactly); then each counted pair vector d = x_j - x_a carries
    (1/S_a) [sigma' <H_a, Y(d) - q_a> d/r + sigma grad_d <H_a, Y(d)>] + sigma' <K_a, q_j - LQ_a> d/r.
    The reference's own force (steinhardt_Local_dcv_AVE_ij/jk_kernel :1623-2130) is not a
    pin: it moves group atoms only."""
    st = _lq_state(x_all, off, idx, n, own, l, swr, eps, group)
    cv, Sf, N = _lq_reduce(st, swcv, mode)
    g, S, q, live, nb, LQ, ql, cm = (st[k] for k in ("g", "S", "q", "live", "nb", "LQ", "ql", "cm"))
    K = np.zeros((n, l + 1), complex)
    for i in range(n):
        if not (g[i] and live[i] and ql[i] > 0.0):
            continue
        f, fp = sw_f(swcv, ql[i]), sw_df(swcv, ql[i])
        if mode == "MEAN_SOLID":
            w = (f + fp * (ql[i] - cv)) / Sf if Sf != 0.0 else 0.0
        elif mode == "NSOLID":
            w = fp
        else:
            w = fp / N if N > 0 else 0.0
        K[i] = w * 4.0 * np.pi / (2 * l + 1) * cm * LQ[i] / ql[i] / (S[i] + 1.0)
    H = K.copy()
    for i in range(n):
        for (o, s, _, _, _) in nb[i]:
            H[o] += s * K[i]

    def dot(A, B):
        return float(np.sum(A.real * B.real + A.imag * B.imag))

    out = np.zeros((n, 3))
    for a in range(n):
        if not live[a]:
            continue
        for (o, s, ds, d, r) in nb[a]:
            Y = ylm(l, d)
            dY = ylm_grad(l, d)                      # (l+1, 3) complex
            gvec = (ds * dot(H[a], Y - q[a]) * d / r + s * (H[a].real @ dY.real + H[a].imag @ dY.imag)) / S[a]
            gvec = gvec + ds * dot(K[a], q[o] - LQ[a]) * d / r
            out[a] -= gvec
            out[o] += gvec
    return out
