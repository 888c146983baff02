"""The batched complex LU solve (mm_zgesv_batched) against LAPACK (numpy) on
seeded random systems of several sizes and right-hand-side counts, pivoting
forced by a zero leading entry, and the info / pivot-range outputs."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solve(A, B):
    from molmot_amd.hip import _lib
    dev = torch.device("cuda", 0)
    At = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
    Bt = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    nb, n, m = B.shape
    info = torch.zeros(nb, dtype=torch.int32, device=dev)
    umin = torch.empty(nb, dtype=torch.float64, device=dev)
    umax = torch.empty(nb, dtype=torch.float64, device=dev)
    rc = _lib.lib().mm_zgesv_batched(n, m, nb, _lib.ptr(At), _lib.ptr(Bt), _lib.ptr(info), _lib.ptr(umin),
                                     _lib.ptr(umax), _lib.stream_handle())
    _lib.check(rc, "mm_zgesv_batched")
    torch.cuda.synchronize()
    return Bt.cpu().numpy(), info.cpu().numpy(), umin.cpu().numpy(), umax.cpu().numpy()


@pytest.mark.parametrize("n,m", [(1, 1), (7, 3), (64, 1), (256, 1), (256, 256), (300, 5), (96, 700)])
def test_zgesv_vs_lapack(n, m):
    rng = np.random.default_rng(n * 1000 + m)
    nb = 4
    A = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    if n > 1:
        A[:, 0, 0] = 0.0                   # the first pivot must come from another row
    B = rng.normal(size=(nb, n, m)) + 1j * rng.normal(size=(nb, n, m))
    print(f"zgesv n={n} nrhs={m}", flush=True)
    X, info, umin, umax = _solve(A, B)
    Xr = np.linalg.solve(A, B)
    assert (info == 0).all() and (umin > 0).all() and (umax >= umin).all()
    assert np.max(np.abs(X - Xr)) < 1e-9 * np.max(np.abs(Xr))


def test_zgesv_singular_reports_info():
    rng = np.random.default_rng(3)
    n = 40
    A = rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n))
    A[1, :, 17] = 0.0                     # column 17 zero: elimination meets a zero pivot there
    B = rng.normal(size=(2, n, 1)) + 0j
    _, info, umin, _ = _solve(A, B)
    assert info[0] == 0 and info[1] == 18 and umin[1] == 0.0
