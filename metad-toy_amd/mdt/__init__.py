"""Python handle on the MI355X metadynamics library (lib/libmdt.so, C ABI in include/mdt.h).

The library is the product; this module only marshals numpy arrays through
ctypes so tests and the benchmark can drive it.  It never computes a CV
itself: every number comes from the gfx950 kernels, and loading fails loudly
when the built library is missing.

Interface mirrored: the LAMMPS ``fix metad`` of the Metad-toy plugin
(src/core/fix_crystallize.cu:32-472): ``Fix(ctx, "GAUSSIAN 0.01 0.005 10 PACE 10 CAL NAME Q6 STEINH Q 6 all ...")``
takes the same keyword list, ``post_force`` is one MD step, ``scalar``/``vector``
are compute_scalar/compute_vector, ``extract`` is Fix::extract("colvar_peratom:...").
"""
import ctypes as C
import os
import shlex

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmdt.so")

MDT_OK = 0
_lib = None


class MdtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdt error {code}: {msg}")
        self.code = code


def lib():
    """Load lib/libmdt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C metad-toy_amd)")
    L = C.CDLL(LIB_PATH)
    P, I, I64, D, U64 = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_uint64
    dp, ip, u8p = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
    sig = {
        "mdt_version": (C.c_char_p, []),
        "mdt_create": (I, [C.POINTER(P), I, I]),
        "mdt_destroy": (None, [P]),
        "mdt_last_error": (C.c_char_p, [P]),
        "mdt_set_box": (I, [P, I, dp, dp, ip]),
        "mdt_bind_positions": (I, [P, I, P, I64, I64, I]),
        "mdt_bind_neighbors": (I, [P, I, P, P, I64, P, I]),
        "mdt_bind_group": (I, [P, I, u8p]),
        "mdt_bind_group_named": (I, [P, I, C.c_char_p, u8p]),
        "mdt_bind_share": (I, [P, I, u8p]),
        "mdt_bind_types": (I, [P, I, ip]),
        "mdt_fix_partials": (I, [P, I64, dp, I64, C.POINTER(I64)]),
        "mdt_fix_resume": (I, [P, I64, dp, P, I]),
        "mdt_bind_system": (I, [P, I, P]),
        "mdt_update_positions": (I, [P, I, P, I]),
        "mdt_attach_positions": (I, [P, I, P]),
        "mdt_fix_create": (I, [P, I, C.POINTER(C.c_char_p), C.POINTER(P)]),
        "mdt_fix_destroy": (None, [P]),
        "mdt_fix_init": (I, [P]),
        "mdt_fix_param": (I, [P, C.c_char_p, dp, I, C.POINTER(I)]),
        "mdt_fix_post_force": (I, [P, I64, P, I]),
        "mdt_fix_run": (I, [P, I64, I64]),
        "mdt_fix_run_steps": (I, [P, I64, I64, I]),
        "mdt_fix_sync": (I, [P]),
        "mdt_fix_profile": (I, [P, I64, I64, C.c_char_p, I, dp, I, C.POINTER(I)]),
        "mdt_fix_scalar": (I, [P, I, dp]),
        "mdt_fix_vector": (I, [P, I, I, dp]),
        "mdt_fix_cv_value": (I, [P, I, C.c_char_p, dp]),
        "mdt_fix_cv_weight": (I, [P, I, C.c_char_p, dp]),
        "mdt_fix_extract_peratom": (I, [P, I, C.c_char_p, dp]),
        "mdt_fix_selection": (I, [P, I, C.c_char_p, ip, ip, ip]),
        "mdt_fix_get_forces": (I, [P, I, dp]),
        "mdt_fix_zero_forces": (I, [P]),
        "mdt_fix_virial": (I, [P, I, ip, dp]),
        "mdt_fix_virial_noalloc": (I, [P, I, P, P]),
        "mdt_fix_mlp_refit": (I, [P, I]),
        "mdt_fix_mlp_status": (I, [P, I, dp, ip, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64)]),
        "mdt_fix_mlp_params": (I, [P, I, I, C.POINTER(C.c_float), dp]),
        "mdt_fix_bias_eval": (I, [P, I, I64, dp, dp, dp]),
        "mdt_fix_bias_add_hills": (I, [P, I, I64, dp, dp, D]),
        "mdt_fix_grid": (I, [P, I, dp, I64]),
        "mdt_fix_hills": (I, [P, I, C.POINTER(I64), dp, I64]),
        "mdt_mlp_forward": (I, [P, C.POINTER(C.c_float), dp, I64, dp, dp]),
        "mdt_mlp_loss_grad": (I, [P, C.POINTER(C.c_float), dp, I64, dp, dp, C.POINTER(C.c_float)]),
        "mdt_fes_sum": (I, [P, I, I64, dp, dp, I, dp, I, dp, dp, dp]),
        "mdt_neighbor_build": (I, [P, P, I64, dp, dp, D, I, C.POINTER(I64), C.POINTER(I64)]),
        "mdt_neighbor_fetch": (I, [P, dp, ip, ip, ip]),
        "mdt_neighbor_bind": (I, [P, I]),
        "mdt_fes_marginal": (I, [P, dp, I, I, I, D, D, dp]),
        "mdt_system_create": (I, [C.POINTER(P)]),
        "mdt_system_destroy": (None, [P]),
        "mdt_system_error": (C.c_char_p, [P]),
        "mdt_system_set_box": (I, [P, dp, dp]),
        "mdt_system_get_box": (None, [P, dp, dp]),
        "mdt_system_add_fcc": (I, [P, I, I, I, D, D, D, D, U64]),
        "mdt_system_add_uniform": (I, [P, I64, D, D, U64]),
        "mdt_system_set_positions": (I, [P, dp, I64]),
        "mdt_system_build": (I, [P, D]),
        "mdt_system_n_local": (I64, [P]),
        "mdt_system_n_all": (I64, [P]),
        "mdt_system_nnz": (I64, [P]),
        "mdt_system_copy": (I, [P, dp, ip, ip, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


class System:
    """Host-built synthetic system + LAMMPS-style full neighbour list."""

    def __init__(self):
        self._h = C.c_void_p()
        if lib().mdt_system_create(C.byref(self._h)) != 0:
            raise MdtError(-1, "mdt_system_create failed")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mdt_system_destroy(self._h)
            self._h = None

    def _chk(self, rc):
        if rc != 0:
            raise MdtError(rc, lib().mdt_system_error(self._h).decode())

    def set_box(self, lo, hi):
        self._chk(lib().mdt_system_set_box(self._h, _dp(np.asarray(lo, float)), _dp(np.asarray(hi, float))))
        return self

    def box(self):
        lo, hi = np.zeros(3), np.zeros(3)
        lib().mdt_system_get_box(self._h, _dp(lo), _dp(hi))
        return lo, hi

    def add_fcc(self, nx, ny, nz, a, sigma_below=0.0, sigma_above=0.0, z_split=0.0, seed=1):
        self._chk(lib().mdt_system_add_fcc(self._h, nx, ny, nz, a, sigma_below, sigma_above, z_split, seed))
        return self

    def add_uniform(self, n, z0, z1, seed=2):
        self._chk(lib().mdt_system_add_uniform(self._h, n, z0, z1, seed))
        return self

    def set_positions(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self._chk(lib().mdt_system_set_positions(self._h, _dp(x), x.shape[0]))
        return self

    def build(self, list_cutoff):
        self._chk(lib().mdt_system_build(self._h, list_cutoff))
        return self

    @property
    def n_local(self):
        return lib().mdt_system_n_local(self._h)

    @property
    def n_all(self):
        return lib().mdt_system_n_all(self._h)

    def arrays(self):
        """x_all (n_all,3), offsets (n_local+1), idx (nnz), ghost_owner (n_all-n_local)."""
        L = lib()
        nl, na, nz = L.mdt_system_n_local(self._h), L.mdt_system_n_all(self._h), L.mdt_system_nnz(self._h)
        na = max(na, nl)
        x = np.zeros((na, 3))
        off = np.zeros(nl + 1, np.int32)
        idx = np.zeros(max(nz, 1), np.int32)
        own = np.zeros(max(na - nl, 1), np.int32)
        self._chk(L.mdt_system_copy(self._h, _dp(x), _ip(off), _ip(idx), _ip(own)))
        return x, off, idx[:nz], own[: na - nl]


class Context:
    """One HIP device (device=-1: parse-only context without a GPU)."""

    def __init__(self, device=0, n_walkers=1):
        self._h = C.c_void_p()
        rc = lib().mdt_create(C.byref(self._h), device, n_walkers)
        if rc != 0:
            raise MdtError(rc, "mdt_create failed (no HIP device?)")
        self.n_walkers = n_walkers

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mdt_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise MdtError(rc, lib().mdt_last_error(self._h).decode())

    def set_box(self, walker, lo, hi):
        per = (C.c_int32 * 3)(1, 1, 1)
        self._chk(lib().mdt_set_box(self._h, walker, _dp(np.asarray(lo, float)), _dp(np.asarray(hi, float)),
                                    C.cast(per, C.POINTER(C.c_int32))))

    def bind_positions(self, walker, x_all, n_local):
        x = np.ascontiguousarray(x_all, np.float64)
        self._chk(lib().mdt_bind_positions(self._h, walker, x.ctypes.data, n_local, x.shape[0], 0))

    def bind_neighbors(self, walker, offsets, idx, ghost_owner):
        off = np.ascontiguousarray(offsets, np.int32)
        ix = np.ascontiguousarray(idx, np.int32)
        ow = np.ascontiguousarray(ghost_owner, np.int32)
        self._chk(lib().mdt_bind_neighbors(self._h, walker, off.ctypes.data, ix.ctypes.data, off.size - 1,
                                           ow.ctypes.data, 0))

    def bind_group(self, walker, mask):
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        self._chk(lib().mdt_bind_group(self._h, walker,
                                       None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8))))

    def bind_group_named(self, walker, name, mask):
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        self._chk(lib().mdt_bind_group_named(self._h, walker, name.encode(),
                                             None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8))))

    def bind_types(self, walker, types):
        t = None if types is None else np.ascontiguousarray(types, np.int32)
        self._chk(lib().mdt_bind_types(self._h, walker, _ip(t)))

    def bind_share(self, walker, mine):
        m = None if mine is None else np.ascontiguousarray(mine, np.uint8)
        self._chk(lib().mdt_bind_share(self._h, walker, None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8))))

    def bind_system(self, walker, system):
        self._chk(lib().mdt_bind_system(self._h, walker, system._h))

    def build_neighbors(self, x_local, lo, hi, list_cutoff):
        """Ghost images and full lists built on this context's device (mdt_neighbor_build);
        returns (n_all, nnz).  fetch_neighbors copies them out, bind_built_neighbors binds them."""
        x = np.ascontiguousarray(x_local, np.float64).reshape(-1, 3)
        na, nz = C.c_int64(), C.c_int64()
        self._chk(lib().mdt_neighbor_build(self._h, x.ctypes.data, x.shape[0], _dp(np.asarray(lo, float)),
                                           _dp(np.asarray(hi, float)), list_cutoff, 0, C.byref(na), C.byref(nz)))
        self._nl = (x.shape[0], na.value, nz.value)
        return na.value, nz.value

    def fetch_neighbors(self):
        """x_all (n_all,3), offsets (n+1), idx (nnz), ghost_owner (n_all-n) of the last device build."""
        n, na, nz = self._nl
        x = np.zeros((na, 3))
        off = np.zeros(n + 1, np.int32)
        idx = np.zeros(max(nz, 1), np.int32)
        own = np.zeros(max(na - n, 1), np.int32)
        self._chk(lib().mdt_neighbor_fetch(self._h, _dp(x), _ip(off), _ip(idx), _ip(own)))
        return x, off, idx[:nz], own[: na - n]

    def bind_built_neighbors(self, walker):
        self._chk(lib().mdt_neighbor_bind(self._h, walker))

    def update_positions(self, walker, x_all):
        x = np.ascontiguousarray(x_all, np.float64)
        self._chk(lib().mdt_update_positions(self._h, walker, x.ctypes.data, 0))

    def update_positions_device(self, walker, x_dev):
        """mdt_update_positions from a device address (n_all*3 doubles)."""
        self._chk(lib().mdt_update_positions(self._h, walker, x_dev, 1))

    def attach_positions(self, walker, x_dev):
        """mdt_attach_positions: every step reads the walker's positions from the device
        address x_dev (None detaches)."""
        self._chk(lib().mdt_attach_positions(self._h, walker, x_dev))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def fes_sum(ctx, hills, x, y=None, lo=None, d=None):
    """Sum of a HILLS list's Gaussians on a grid (mdt_fes_sum): hills rows (step, s..., h, sigma);
    x (and y, lo[2], d[2] for two dimensions) as tools/plot_hills.py builds them.  Returns F of
    shape (len(x),) or (len(x), len(y))."""
    hills = np.ascontiguousarray(np.atleast_2d(hills), np.float64)
    D = hills.shape[1] - 3
    x = np.ascontiguousarray(x, np.float64)
    if D == 2:
        y = np.ascontiguousarray(y, np.float64)
        lo, d = np.ascontiguousarray(lo, np.float64), np.ascontiguousarray(d, np.float64)
        out = np.zeros((x.size, y.size))
        ctx._chk(lib().mdt_fes_sum(ctx._h, 2, hills.shape[0], _dp(hills), _dp(x), x.size, _dp(y), y.size,
                                   _dp(lo), _dp(d), _dp(out)))
    else:
        out = np.zeros(x.size)
        ctx._chk(lib().mdt_fes_sum(ctx._h, D, hills.shape[0], _dp(hills), _dp(x), x.size, None, 0, None, None,
                                   _dp(out)))
    return out


def fes_marginal(ctx, F, axis, kT, bin_size):
    """Marginal FES of a 2-D F along axis (1: F(CV1), 0: F(CV2)), minimum shifted to 0."""
    F = np.ascontiguousarray(F, np.float64)
    n0, n1 = F.shape
    out = np.zeros(n0 if axis == 1 else n1)
    ctx._chk(lib().mdt_fes_marginal(ctx._h, _dp(F), n0, n1, axis, kT, bin_size, _dp(out)))
    return out


def mlp_forward(ctx, params, norm, s):
    """Network V at points s (n,3) on the MFMA inference kernel."""
    p = np.ascontiguousarray(params, np.float32)
    nm = np.ascontiguousarray(norm, np.float64)
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    V = np.zeros(s.shape[0])
    ctx._chk(lib().mdt_mlp_forward(ctx._h, _fp(p), _dp(nm), s.shape[0], _dp(s), _dp(V)))
    return V


def mlp_loss_grad(ctx, params, norm, s, v):
    """Gradient of (1/n) sum (y - v/out_scale)^2 over the 4481 parameters (MFMA training kernel)."""
    p = np.ascontiguousarray(params, np.float32)
    nm = np.ascontiguousarray(norm, np.float64)
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, np.float64)
    g = np.zeros(4481, np.float32)
    ctx._chk(lib().mdt_mlp_loss_grad(ctx._h, _fp(p), _dp(nm), s.shape[0], _dp(s), _dp(v), _fp(g)))
    return g


class Fix:
    """``fix metad`` on every walker of a context."""

    def __init__(self, ctx, args):
        if isinstance(args, str):
            args = shlex.split(args)
        self.ctx = ctx
        self.args = list(args)
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        self._h = C.c_void_p()
        ctx._chk(lib().mdt_fix_create(ctx._h, len(args), arr, C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mdt_fix_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        self.ctx._chk(rc)

    def init(self):
        self._chk(lib().mdt_fix_init(self._h))

    def param(self, key):
        """Parsed values of a fix keyword (mdt_fix_param): GAUSSIAN, PACE, CV_dim, TEMP, BOLTZ,
        WT, METAD_RESTART, Gaussian_Hill_type, SIGMAS, DIM (lower, upper, bins per dimension)."""
        n = C.c_int()
        self._chk(lib().mdt_fix_param(self._h, key.encode(), None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        self._chk(lib().mdt_fix_param(self._h, key.encode(), _dp(out), n.value, C.byref(n)))
        return out[: n.value]

    def post_force(self, step, f=None):
        """One MD step; f (host float64 array, n_local_total x 3) receives += bias forces."""
        if f is None:
            self._chk(lib().mdt_fix_post_force(self._h, step, None, 0))
        else:
            assert f.dtype == np.float64 and f.flags.c_contiguous
            self._chk(lib().mdt_fix_post_force(self._h, step, f.ctypes.data, 0))

    def partials(self, step):
        """mdt_fix_partials: this context's reduction vector of the step (all walkers)."""
        n = C.c_int64()
        self._chk(lib().mdt_fix_partials(self._h, step, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        self._chk(lib().mdt_fix_partials(self._h, step, _dp(out), n.value, C.byref(n)))
        return out[: n.value]

    def resume(self, step, reduced, f=None):
        """mdt_fix_resume with the reduction vector summed over the contexts."""
        r = np.ascontiguousarray(reduced, np.float64)
        if f is None:
            self._chk(lib().mdt_fix_resume(self._h, step, _dp(r), None, 0))
        else:
            assert f.dtype == np.float64 and f.flags.c_contiguous
            self._chk(lib().mdt_fix_resume(self._h, step, _dp(r), f.ctypes.data, 0))

    def run(self, step0, nsteps, per_launch=None):
        """nsteps steps on the device; per_launch = captured steps per graph launch
        (default: the library's batch), 0 for plain kernel launches."""
        if per_launch is None:
            self._chk(lib().mdt_fix_run(self._h, step0, nsteps))
        else:
            self._chk(lib().mdt_fix_run_steps(self._h, step0, nsteps, per_launch))

    def sync(self):
        self._chk(lib().mdt_fix_sync(self._h))

    def profile(self, step0, nsteps):
        """{kernel name: mean ms per launch} over nsteps plain-launch steps (HIP events)."""
        names = C.create_string_buffer(4096)
        ms = np.zeros(64)
        nk = C.c_int()
        self._chk(lib().mdt_fix_profile(self._h, step0, nsteps, names, 4096, _dp(ms), 64, C.byref(nk)))
        keys = names.value.decode().split(",") if nk.value else []
        return {k: float(ms[i]) for i, k in enumerate(keys)}

    def scalar(self, walker=0):
        v = C.c_double()
        self._chk(lib().mdt_fix_scalar(self._h, walker, C.byref(v)))
        return v.value

    def vector(self, n, walker=0):
        v = C.c_double()
        self._chk(lib().mdt_fix_vector(self._h, walker, n, C.byref(v)))
        return v.value

    def cv(self, cal_func, walker=0):
        v = C.c_double()
        self._chk(lib().mdt_fix_cv_value(self._h, walker, cal_func.encode(), C.byref(v)))
        return v.value

    def cv_weight(self, cal_func, walker=0):
        v = C.c_double()
        self._chk(lib().mdt_fix_cv_weight(self._h, walker, cal_func.encode(), C.byref(v)))
        return v.value

    def extract(self, key, n_local, width=1, walker=0):
        out = np.zeros(n_local * width)
        self._chk(lib().mdt_fix_extract_peratom(self._h, walker, key.encode(), _dp(out)))
        return out if width == 1 else out.reshape(n_local, width)

    def selection(self, cal, n_local, N, walker=0):
        sel = np.zeros(n_local * N, np.int32)
        cnt = np.zeros(n_local, np.int32)
        ns = np.zeros(n_local, np.int32)
        self._chk(lib().mdt_fix_selection(self._h, walker, cal.encode(), _ip(sel), _ip(cnt), _ip(ns)))
        return sel.reshape(n_local, N), cnt, ns

    def forces(self, n_local, walker=0):
        f = np.zeros((n_local, 3))
        self._chk(lib().mdt_fix_get_forces(self._h, walker, _dp(f)))
        return f

    def virial(self, walker=0, image=None):
        v = np.zeros(6)
        img = None if image is None else np.ascontiguousarray(image, np.int32)
        self._chk(lib().mdt_fix_virial(self._h, walker, _ip(img), _dp(v)))
        return v

    def virial_noalloc(self, walker, image_dev, v_dev):
        """mdt_fix_virial_noalloc: image_dev (device int32, 3 per local atom, or None) and
        v_dev (device float64, 6) are raw device addresses; nothing is allocated."""
        self._chk(lib().mdt_fix_virial_noalloc(self._h, walker, image_dev, v_dev))

    def mlp_refit(self, walker=0):
        self._chk(lib().mdt_fix_mlp_refit(self._h, walker))

    def mlp_status(self, walker=0):
        err = C.c_double()
        ok = C.c_int32()
        fits, fb, nfit = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(lib().mdt_fix_mlp_status(self._h, walker, C.byref(err), C.byref(ok), C.byref(fits),
                                           C.byref(fb), C.byref(nfit)))
        return {"err": err.value, "ok": ok.value, "fits": fits.value, "fallbacks": fb.value, "n_fit": nfit.value}

    def mlp_params(self, walker=0, live=True):
        p = np.zeros(4481, np.float32)
        norm = np.zeros(8)
        self._chk(lib().mdt_fix_mlp_params(self._h, walker, 1 if live else 0,
                                           p.ctypes.data_as(C.POINTER(C.c_float)), _dp(norm)))
        return p, norm

    def zero_forces(self):
        self._chk(lib().mdt_fix_zero_forces(self._h))

    def bias_eval(self, s, walker=0):
        s = np.ascontiguousarray(np.atleast_2d(s), np.float64)
        n, D = s.shape
        V = np.zeros(n)
        dV = np.zeros((n, D))
        self._chk(lib().mdt_fix_bias_eval(self._h, walker, n, _dp(s), _dp(V), _dp(dV)))
        return V, dV

    def add_hills(self, s, h, sigma, walker=0):
        s = np.ascontiguousarray(np.atleast_2d(s), np.float64)
        h = np.ascontiguousarray(h, np.float64)
        self._chk(lib().mdt_fix_bias_add_hills(self._h, walker, h.size, _dp(s), _dp(h), sigma))

    def grid(self, n, walker=0):
        g = np.zeros(n)
        self._chk(lib().mdt_fix_grid(self._h, walker, _dp(g), n))
        return g

    def hills(self, D, walker=0, cap=1 << 20):
        nh = C.c_int64()
        self._chk(lib().mdt_fix_hills(self._h, walker, C.byref(nh), None, 0))
        rec = np.zeros(max(nh.value, 1) * (D + 3))
        self._chk(lib().mdt_fix_hills(self._h, walker, C.byref(nh), _dp(rec), nh.value))
        return rec[: nh.value * (D + 3)].reshape(nh.value, D + 3)
