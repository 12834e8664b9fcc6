"""mdt_fix_virial_noalloc, the entry point that allocates nothing, against mdt_fix_virial.

Two walkers of different sizes share one context (so one device scratch array): the
allocation-free virial of the larger walker, then of the smaller one, each equal bit for
bit to the allocating entry point's, with image flags and without.  The test's own device buffers
(image flags, result) come from the HIP runtime the library itself links.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import fcc_system, q46_fix

pytestmark = pytest.mark.gpu


class _Dev:
    """A device buffer from the library's HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def test_virial_noalloc_matches_virial(mdt):
    systems = [fcc_system(mdt, 10, sigma=0.15, seed=71), fcc_system(mdt, 7, sigma=0.15, seed=72)]
    ns = [s.n_local for s in systems]
    print(f"[no-alloc virial] walkers n={ns}", flush=True)
    ctx = mdt.Context(0, 2)
    for w, s in enumerate(systems):
        ctx.bind_system(w, s)
    fix = mdt.Fix(ctx, q46_fix(pace=2, wt=1))
    fix.zero_forces()
    for step in range(3):
        fix.post_force(step)            # bias forces into the context's device buffer
    rng = np.random.default_rng(9)
    for w in (0, 1):                    # the second call has the smaller input
        img = rng.integers(-2, 3, size=(ns[w], 3)).astype(np.int32)
        ref = fix.virial(w, img)
        ref0 = fix.virial(w, None)
        assert np.abs(ref).max() > 0.0
        img_d, v_d = _Dev(img.nbytes), _Dev(6 * 8)
        img_d.put(img)
        v_d.put(np.full(6, np.nan))
        fix.virial_noalloc(w, img_d.p.value, v_d.p.value)
        fix.sync()
        v = v_d.get(np.zeros(6))
        assert np.array_equal(v, ref), (w, v, ref)
        v_d.put(np.full(6, np.nan))
        fix.virial_noalloc(w, None, v_d.p.value)
        fix.sync()
        assert np.array_equal(v_d.get(np.zeros(6)), ref0), w
        img_d.free()
        v_d.free()
    with pytest.raises(mdt.MdtError):
        fix.virial_noalloc(0, None, None)
    fix.close()
    ctx.close()
