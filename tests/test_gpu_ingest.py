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
    for mode in ("host", "device", "attached", "attached_graph"):
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
            if mode == "attached_graph":
                fix.run(step, 1, per_launch=1)
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
    for mode in ("device", "attached", "attached_graph"):
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
