"""Per-step position ingestion (run on an MI355X with -m gpu).

mdt_attach_positions makes every step read the positions from the MD engine's device
buffer (packed at the start of the step, inside its graph); mdt_update_positions with a
device pointer packs the same way.  Both must give the bits of positions bound from the
host, step after step, while the buffer's contents change between steps.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import Mover, fcc_system, q46_fix

pytestmark = pytest.mark.gpu


class _Dev:
    def __init__(self, a):
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.p = C.c_void_p()
        self.n = a.nbytes
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(a.nbytes)) == 0
        self.put(a)

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def free(self):
        self.hip.hipFree(self.p)


def test_attached_and_device_positions_match_host_positions(mdt):
    s = fcc_system(mdt, 8, sigma=0.2, seed=17)
    x, off, idx, own = s.arrays()
    n = s.n_local
    rng = np.random.default_rng(3)
    mv = Mover(x, n, own)
    xl = x[:n].copy()
    traj = []
    for _ in range(6):
        traj.append(mv.set_local(xl).copy())
        xl = xl + rng.normal(scale=0.01, size=xl.shape)
    print(f"[ingest] n={n} steps={len(traj)}", flush=True)
    out = {}
    for mode in ("host", "device", "attached", "attached_graph", "attached_plain"):
        ctx = mdt.Context(0, 1)
        ctx.bind_system(0, s)
        fix = mdt.Fix(ctx, q46_fix(pace=2, wt=1))
        buf = _Dev(traj[0])
        if mode.startswith("attached"):
            ctx.attach_positions(0, buf.p.value)
        fs, cvs = [], []
        for step, xa in enumerate(traj):
            fix.zero_forces()
            if mode == "host":
                ctx.update_positions(0, xa)
            else:
                buf.put(xa)
                if mode == "device":
                    ctx.update_positions_device(0, buf.p.value)
            if mode in ("attached_graph", "attached_plain"):
                fix.run(step, 1, per_launch=1 if mode == "attached_graph" else 0)
                fix.sync()
            else:
                fix.post_force(step)
            fs.append(fix.forces(n))
            cvs.append((fix.cv("Q6.AVE"), fix.cv("Q4.AVE"), fix.scalar()))
        out[mode] = (np.array(fs), cvs)
        buf.free()
        fix.close()
        ctx.close()
    f0, c0 = out["host"]
    assert np.abs(f0).max() > 0.0
    for mode in ("device", "attached", "attached_graph", "attached_plain"):
        f1, c1 = out[mode]
        assert np.array_equal(f0, f1), mode
        assert c0 == c1, mode


def test_attached_positions_drive_near_list_rebuilds(mdt, orc):
    """Lists of ~240 entries (the reference's Q4 deck geometry) go through near lists that are
    rebuilt inside the step whenever an atom has moved far enough; with the positions read
    from an attached buffer inside replayed graphs, every step gives the bits of host-bound
    positions stepped one at a time, and Q4.AVE follows the oracle while the atoms move well
    past the rebuild threshold."""
    s = fcc_system(mdt, 8, a=3.26, sigma=0.05, seed=48, rg=8.0)
    x, off, idx, own = s.arrays()
    n = s.n_local
    rng = np.random.default_rng(9)
    mv = Mover(x, n, own)
    xl = x[:n].copy()
    traj = []
    for _ in range(8):
        traj.append(mv.set_local(xl).copy())
        xl = xl + rng.normal(scale=0.06, size=xl.shape)
    deck = ("GAUSSIAN 0.01 0.005 10 PACE 3 TEMP 300 HILLS_FILE NONE "
            "CAL NAME Q4 STEINH Q 4 all cutoff_r 6.0 cutoff_Natoms 12 "
            "CV_dim 1 SYMBOL v1 Q4.AVE DIM 1 0 1 1000 v1 WT 1")
    print(f"[ingest near] n={n} mean list {off[-1] / n:.0f}", flush=True)
    out = {}
    for mode in ("host", "attached_graph"):
        ctx = mdt.Context(0, 1)
        ctx.bind_system(0, s)
        fix = mdt.Fix(ctx, deck)
        buf = _Dev(traj[0])
        if mode == "attached_graph":
            ctx.attach_positions(0, buf.p.value)
        fs, cvs = [], []
        for step, xa in enumerate(traj):
            fix.zero_forces()
            if mode == "host":
                ctx.update_positions(0, xa)
                fix.post_force(step)
            else:
                buf.put(xa)
                fix.run(step, 1, per_launch=1)
                fix.sync()
            fs.append(fix.forces(n))
            cvs.append((fix.cv("Q4.AVE"), fix.scalar()))
        out[mode] = (np.array(fs), cvs)
        buf.free()
        fix.close()
        ctx.close()
    f0, c0 = out["host"]
    f1, c1 = out["attached_graph"]
    assert np.abs(f0).max() > 0.0
    assert np.array_equal(f0, f1)
    assert c0 == c1
    for step in (0, len(traj) - 1):
        q4 = orc.Env(traj[step], off, idx, own, n, 6.0, 12).q_ave(4)
        assert abs(c0[step][0] - q4) <= 1e-5 * q4, step
    moved = np.abs(traj[-1][:n] - traj[0][:n]).max()
    assert moved > 0.2, moved            # far past half the near-list margin (rc / 128)


def test_attached_positions_exact_selection_paths(mdt, orc):
    """A one-walker context reading an attached engine buffer packs only the FP32
    positions, and its exact FP64 tests read the engine buffer in place.  On a perfect
    lattice with the cutoff at the nearest-neighbour distance every candidate goes
    through the FP64 band test, and with N = 4 the cap binds for most atoms (the
    reference insertion, with exact ties on the unmoved lattice): forces, CVs, the kept sets and the
    virial (which refreshes the FP64 copy first) equal those of host-bound positions
    bit for bit, step after step, and the kept sets follow the oracle."""
    a = 5.26
    rc = a / np.sqrt(2.0)
    s = fcc_system(mdt, 5, a=a, sigma=0.0, seed=4)
    x, off, idx, own = s.arrays()
    n = s.n_local
    mv = Mover(x, n, own)
    rng = np.random.default_rng(8)
    traj = [x.copy()]
    xl = x[:n].copy()
    for _ in range(2):
        xl = xl + rng.normal(scale=1e-4, size=xl.shape)
        traj.append(mv.set_local(xl).copy())
    N = 4
    fixargs = q46_fix(pace=1, wt=1, rc=repr(float(rc)), N=N)
    print(f"[ingest exact] n={n} rc={rc!r} N={N} steps={len(traj)}", flush=True)
    out = {}
    for mode in ("host", "attached"):
        ctx = mdt.Context(0, 1)
        ctx.bind_system(0, s)
        fix = mdt.Fix(ctx, fixargs)
        buf = _Dev(traj[0])
        if mode == "attached":
            ctx.attach_positions(0, buf.p.value)
        res = []
        for step, xa in enumerate(traj):
            fix.zero_forces()
            if mode == "host":
                ctx.update_positions(0, xa)
            else:
                buf.put(xa)
            fix.post_force(step)
            kept = fix.extract("Q6.kept", n, N).astype(np.int64)
            sel, cnt, ns = fix.selection("Q6", n, N)
            res.append((fix.forces(n), (fix.cv("Q6.AVE"), fix.cv("Q4.AVE"), fix.scalar()), sel, cnt,
                        fix.virial(), kept))
        out[mode] = res
        buf.free()
        fix.close()
        ctx.close()
    for (f0, c0, s0, n0, v0, k0), (f1, c1, s1, n1, v1, k1) in zip(out["host"], out["attached"]):
        assert np.abs(f0).max() > 0.0
        assert np.array_equal(f0, f1)
        assert c0 == c1
        assert np.array_equal(s0, s1) and np.array_equal(n0, n1) and np.array_equal(k0, k1)
        assert np.array_equal(v0, v1)
    # every step's kept set and slot order against the oracle's reference insertion
    for xa, (f1, c1, sel, cnt, v1, kept) in zip(traj, out["attached"]):
        env = orc.Env(xa, off, idx, own, n, float(rc), N)
        want = np.full((n, N), -1, np.int64)
        for i in range(n):
            pos = np.sort(env.sel_pos[i][env.sel_pos[i] >= 0])
            want[i, : pos.size] = idx[off[i] + pos]
        assert np.array_equal(kept, want)
        assert np.array_equal(cnt, env.cnt) and np.array_equal(sel, env.sel_idx)
        assert (env.cnt > N).sum() > n // 2           # the cap binds for most atoms
