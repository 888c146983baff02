"""mm_traj_euler_1d_noalloc (no caller workspace: the queue lives in the
library's file-scope scratch array) against mm_traj_euler_1d_ws with a
caller workspace, through the C ABI: bit-identical finals on two calls in a
row, the second smaller, on the work-queue schedule; a size whose queue
exceeds the scratch array is rejected before any launch."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(L, plan, z0, v0, n_steps, noalloc, schedule, seed):
    from molmot_amd.hip import _lib
    n = z0.numel()
    dev = z0.device
    zf = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    vf = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    na = torch.full((n,), -7, dtype=torch.int32, device=dev)
    s = _lib.stream_handle()
    if noalloc:
        rc = L.mm_traj_euler_1d_noalloc(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), 0.16, plan.mass, 1e-5,
                                        n_steps, 0.015, _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na), 2e-3, seed, 0, 0,
                                        schedule, s)
        _lib.check(rc, "mm_traj_euler_1d_noalloc")
    else:
        nb = int(L.mm_traj_euler_workspace_bytes(n, n_steps))
        ws = _lib.queue_workspace(nb, dev)
        rc = L.mm_traj_euler_1d_ws(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), 0.16, plan.mass, 1e-5, n_steps,
                                   0.015, None, None, _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na), 2e-3, seed, 0, 0,
                                   schedule, _lib.ptr(ws), nb, s)
        _lib.check(rc, "mm_traj_euler_1d_ws")
    torch.cuda.synchronize()
    return zf.cpu().numpy(), vf.cpu().numpy(), na.cpu().numpy()


def test_noalloc_matches_workspace_entry_point(mol):
    from molmot_amd.constants import hbar
    from molmot_amd.hip import HIPMolData, _lib
    from molmot_amd.obe.fields import make_dc_mot_beams
    L = _lib.lib()
    plan = HIPMolData(mol, make_dc_mot_beams(mol, -0.2, 0.7, 1.0), 1.0, hbar=hbar)
    rng = np.random.default_rng(17)
    dev = torch.device("cuda", 0)
    for n, n_steps, seed in ((3000, 2049, 5), (700, 1025, 6)):   # second call smaller
        print(f"noalloc: {n} particles x {n_steps} samples, queue schedule", flush=True)
        z0 = torch.from_numpy(rng.normal(0, 1e-3, n)).to(dev)
        v0 = torch.from_numpy(rng.uniform(-12, 12, n)).to(dev)
        a = _run(L, plan, z0, v0, n_steps, True, 2, seed)
        b = _run(L, plan, z0, v0, n_steps, False, 2, seed)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
        assert (a[2] > 0).all() and np.isfinite(a[0]).all()
        c = _run(L, plan, z0, v0, n_steps, True, 0, seed)      # auto schedule
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[2], c[2])


def test_noalloc_rejects_queues_beyond_the_scratch_array(mol):
    from molmot_amd.constants import hbar
    from molmot_amd.hip import HIPMolData, _lib
    from molmot_amd.obe.fields import make_dc_mot_beams
    L = _lib.lib()
    assert L.mm_traj_euler_noalloc_max_bytes() == 4 << 20
    plan = HIPMolData(mol, make_dc_mot_beams(mol, -0.2, 0.7, 1.0), 1.0, hbar=hbar)
    n, n_steps = 400_000, 30_000          # 6,250 groups x 234 items > 4 MiB
    assert L.mm_traj_euler_workspace_bytes(n, n_steps) > (4 << 20)
    dev = torch.device("cuda", 0)
    z0 = torch.zeros(n, dtype=torch.float64, device=dev)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    na = torch.empty(n, dtype=torch.int32, device=dev)
    rc = L.mm_traj_euler_1d_noalloc(plan.handle, n, _lib.ptr(z0), _lib.ptr(z0), 0.16, plan.mass, 1e-5, n_steps,
                                    0.015, _lib.ptr(out), _lib.ptr(out), _lib.ptr(na), 0.0, 0, 0, 0, 2,
                                    _lib.stream_handle())
    assert rc == -1 and b"scratch" in L.mm_last_error()
