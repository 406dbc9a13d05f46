"""hip_diag_hamil_noalloc (the allocation-free variant of hip_diag_hamil) through the C ABI
against hip_diag_hamil on the same input, on two calls in a row, the second smaller:
eigenvalues and eigenvectors identical to FP64 rounding (same kernels, same order), and
both against LAPACK."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _herm(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return (X + X.conj().T) / 2


def _solve(fn, H):
    import torch
    n = H.shape[0]
    A = torch.from_numpy(H.T.copy()).cuda()         # column-major buffer of H
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = fn(A.data_ptr(), w.data_ptr(), n)
    assert rc == 0, rc
    return w.cpu().numpy(), A.cpu().numpy().T


@pytest.mark.parametrize("sizes", [(256, 37), (200, 1)])
def test_noalloc_matches_hip_diag_hamil(sizes):
    from rlm_amd import lib
    L = lib()
    for k, n in enumerate(sizes):
        print(f"[test_noalloc_matches_hip_diag_hamil] call {k}: n={n}", flush=True)
        H = _herm(n, 100 + n)
        w0, U0 = _solve(L.hip_diag_hamil, H)
        w1, U1 = _solve(L.hip_diag_hamil_noalloc, H)
        nrm = max(np.linalg.norm(H, 2), 1e-300)
        assert np.abs(w1 - w0).max() <= 1e-12 * nrm
        assert np.abs(U1 - U0).max() <= 1e-10
        assert np.abs(w1 - np.linalg.eigvalsh(H)).max() <= 1e-12 * nrm * max(1, n / 64)
        assert np.abs(H @ U1 - U1 * w1[None, :]).max() <= 1e-12 * nrm * max(1, n / 64)


def test_noalloc_rejects_large_and_nonfinite():
    import torch
    from rlm_amd import lib
    print("[test_noalloc_rejects_large_and_nonfinite]", flush=True)
    A = torch.zeros((300, 300), dtype=torch.complex128, device="cuda")
    w = torch.zeros(300, dtype=torch.float64, device="cuda")
    assert lib().hip_diag_hamil_noalloc(A.data_ptr(), w.data_ptr(), 257) == -22
    H = _herm(40, 3)
    H[7, 2] = np.nan
    B = torch.from_numpy(H.T.copy()).cuda()
    assert lib().hip_diag_hamil_noalloc(B.data_ptr(), w.data_ptr(), 40) == 1
