"""GPU parity of the batched Hermitian eigensolver (librlm_hip.so) against LAPACK zheevd
(numpy.linalg.eigh), the routine the reference calls in diag_wrappers/mkl_diag.cpp:8.

Tolerances (FP64): eigenvalues |w - w_ref| <= 1e-12 * ||A||, residual
||A U - U diag(w)|| <= 1e-12 * ||A||, orthogonality ||U^H U - I|| <= 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rand_herm(rng, b, n, scale=1.0):
    X = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
    return scale * (X + X.conj().transpose(0, 2, 1)) / 2


def _check(H, w, U):
    b, n, _ = H.shape
    for k in range(b):
        A = H[k]
        nrm = max(np.linalg.norm(A, 2), 1e-300)
        wr = np.linalg.eigvalsh(A)
        assert np.all(np.diff(w[k]) >= 0), "eigenvalues not ascending"
        assert np.abs(w[k] - wr).max() <= TOL * nrm * max(1, n / 64), (k, np.abs(w[k] - wr).max() / nrm)
        R = A @ U[k] - U[k] * w[k][None, :]
        assert np.abs(R).max() <= TOL * nrm * max(1, n / 64), (k, np.abs(R).max() / nrm)
        O = U[k].conj().T @ U[k] - np.eye(n)
        assert np.abs(O).max() <= TOL * max(1, n / 64), (k, np.abs(O).max())


@pytest.mark.parametrize("n,b", [(1, 3), (2, 5), (3, 4), (7, 3), (16, 8), (33, 4), (64, 3), (72, 6), (128, 2),
                                 (255, 2), (256, 3)])
def test_random_hermitian(n, b):
    import torch
    from rlm_amd import heevd_batched
    print(f"[test_random_hermitian] n={n} b={b}", flush=True)
    rng = np.random.default_rng(100 + n)
    H = _rand_herm(rng, b, n)
    w, U = heevd_batched(torch.from_numpy(H).cuda())
    _check(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("variant", ["fused", "blocked", "lower", "resident", "hybrid"])
@pytest.mark.parametrize("n,b", [(5, 3), (17, 4), (100, 3), (128, 2), (256, 2)])
def test_tridiagonalisation_variants(monkeypatch, variant, n, b):
    """Every tridiagonalisation kernel (RLM_HETRD=fused | blocked | lower | resident | hybrid) against
    LAPACK; the register-resident FP64 kernel covers n <= 128 (beyond, the lower kernel runs)."""
    import torch
    from rlm_amd import heevd_batched
    print(f"[test_tridiagonalisation_variants] {variant} n={n}", flush=True)
    monkeypatch.setenv("RLM_HETRD", variant)
    rng = np.random.default_rng(7 * n + 1)
    H = _rand_herm(rng, b, n)
    w, U = heevd_batched(torch.from_numpy(H).cuda())
    _check(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("n", [384, 512])
def test_large_sizes(n):
    import torch
    from rlm_amd import heevd_batched
    print(f"[test_large_sizes] n={n}", flush=True)
    rng = np.random.default_rng(n)
    H = _rand_herm(rng, 2, n)
    w, U = heevd_batched(torch.from_numpy(H).cuda())
    _check(H, w.cpu().numpy(), U.cpu().numpy())


def test_degenerate_and_structured():
    import torch
    from rlm_amd import heevd_batched
    print("[test_degenerate_and_structured]", flush=True)
    n = 40
    rng = np.random.default_rng(7)
    mats = []
    mats.append(np.eye(n, dtype=np.complex128))                      # fully degenerate
    mats.append(np.zeros((n, n), dtype=np.complex128))               # zero matrix
    mats.append(np.diag(np.arange(n, dtype=float)).astype(np.complex128))   # already diagonal
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    ev = np.repeat([-1.0, 0.5, 2.0, 2.0 + 1e-13], n // 4)
    mats.append(Q @ np.diag(ev) @ Q.conj().T)                        # clusters
    T = np.diag(np.abs(np.arange(n) - n // 2).astype(float)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    mats.append(T.astype(np.complex128))                             # Wilkinson-like (real)
    mats.append(1e-100 * _rand_herm(rng, 1, n)[0])                   # tiny scale
    mats.append(1e100 * _rand_herm(rng, 1, n)[0])                    # huge scale
    H = np.stack([(M + M.conj().T) / 2 for M in mats])
    w, U = heevd_batched(torch.from_numpy(H).cuda())
    _check(H, w.cpu().numpy(), U.cpu().numpy())


def test_uplo_and_hamil_entry():
    """hip_diag_hamil reads only the lower triangle (cusolver_diag.cu:8); hip_heevd reads 'U' (mkl_diag.cpp:6)."""
    import ctypes
    import torch
    from rlm_amd import lib
    print("[test_uplo_and_hamil_entry]", flush=True)
    n = 48
    rng = np.random.default_rng(3)
    H = _rand_herm(rng, 1, n)[0]
    junk = H.copy()
    iu = np.triu_indices(n, 1)
    junk[iu] = 99.0 + 7j                       # garbage in the strict upper triangle
    A = torch.from_numpy(junk.T.copy()).cuda()  # column-major buffer of `junk`
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert lib().hip_diag_hamil(A.data_ptr(), w.data_ptr(), n) == 0
    wr = np.linalg.eigvalsh(H)
    assert np.abs(w.cpu().numpy() - wr).max() <= 1e-12 * np.linalg.norm(H, 2)
    U = A.cpu().numpy().T
    assert np.abs(H @ U - U * w.cpu().numpy()[None, :]).max() <= 1e-12 * np.linalg.norm(H, 2)
    # host entry, upper triangle read
    junk2 = H.copy()
    il = np.tril_indices(n, -1)
    junk2[il] = -5.0
    buf = np.ascontiguousarray(junk2.T)        # column-major host buffer
    wh = np.zeros(n)
    assert lib().hip_heevd(buf.ctypes.data_as(ctypes.c_void_p), wh.ctypes.data_as(ctypes.c_void_p), n) == 0
    assert np.abs(wh - wr).max() <= 1e-12 * np.linalg.norm(H, 2)


def test_diag_hamil_void_entry():
    """hip_diag_hamil_v, the void signature of cusolver_diag_hamil / magma_diag_hamil
    (rlm/main_rlm_gpu.cu:176,216): same result as hip_diag_hamil."""
    import torch
    from rlm_amd import lib
    print("[test_diag_hamil_void_entry]", flush=True)
    n = 40
    H = _rand_herm(np.random.default_rng(11), 1, n)[0]
    A = torch.from_numpy(H.T.copy()).cuda()
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    lib().hip_diag_hamil_v(A.data_ptr(), w.data_ptr(), n)
    torch.cuda.synchronize()
    _check(H[None], w.cpu().numpy()[None], A.cpu().numpy().T[None])


@pytest.mark.parametrize("precision", [0, 1])
def test_nonfinite_input_status(precision):
    """A NaN or an infinity in the triangle read stops the call before any solver kernel
    with status 1 + the index of the first such matrix; d_A and d_W stay untouched, and
    the next call on finite input solves normally."""
    import ctypes
    import torch
    from rlm_amd import lib
    print(f"[test_nonfinite_input_status] precision {precision}", flush=True)
    b, n = 4, 24
    H = _rand_herm(np.random.default_rng(2), b, n)

    def call(Hs, uplo=b"L"):
        A = torch.from_numpy(Hs.transpose(0, 2, 1).copy()).cuda()
        A0 = A.clone()
        W = torch.full((b, n), 7.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        rc = lib().hip_diag_heevd_batched(n, A.data_ptr(), n, n * n, W.data_ptr(), n, b, uplo, precision, None)
        return rc, A, A0, W

    bad = H.copy()
    bad[2, 5, 3] = np.nan                    # lower triangle (row 5, column 3)
    rc, A, A0, W = call(bad)
    assert rc == 3, rc
    assert torch.equal(torch.isnan(A), torch.isnan(A0)) and torch.equal(A.nan_to_num(), A0.nan_to_num())
    assert bool((W == 7.0).all())
    bad = H.copy()
    bad[1, 0, 9] = np.inf + 0j               # upper triangle: ignored by 'L', caught by 'U'
    rc, *_ = call(bad, b"L")
    assert rc == 0
    rc, *_ = call(bad, b"U")
    assert rc == 2, rc
    rc, A, _, W = call(H)
    assert rc == 0
    if precision == 0:
        _check(H, W.cpu().numpy(), A.cpu().numpy().transpose(0, 2, 1))
    # host entry (mkl_heevd's replacement)
    h = np.ascontiguousarray(H[0].T)
    h[0, 0] = np.nan
    wh = np.zeros(n)
    assert lib().hip_heevd(h.ctypes.data_as(ctypes.c_void_p), wh.ctypes.data_as(ctypes.c_void_p), n) == 1


def test_invalid_arguments():
    from rlm_amd import lib
    print("[test_invalid_arguments]", flush=True)
    assert lib().hip_diag_heevd_batched(0, None, 1, 1, None, 1, 1, b"L", 0, None) == -22
    assert lib().hip_diag_heevd_batched(5000, None, 5000, 5000 * 5000, None, 5000, 1, b"L", 0, None) == -22
    assert lib().hip_diag_heevd_batched(600, None, 600, 600 * 600, None, 600, 1, b"L", 0, None) == -22
    assert lib().hip_diag_heevd_batched(4, None, 4, 16, None, 4, 1, b"X", 0, None) == -22
    assert lib().hip_diag_heevd_batched(4, None, 4, 16, None, 4, 0, b"L", 0, None) == 0


def _check_mixed(H, w, U, tol_vec=3e-5):
    """Mixed precision: FP64 Rayleigh-quotient eigenvalues within 1e-9 ||A|| of LAPACK's,
    FP32-accurate eigenvectors (residual and orthogonality within tol_vec * n/64)."""
    b, n, _ = H.shape
    for k in range(b):
        A = H[k]
        nrm = max(np.linalg.norm(A, 2), 1e-300)
        wr = np.linalg.eigvalsh(A)
        sc = max(1, n / 64)
        assert np.abs(np.sort(w[k]) - wr).max() <= 1e-9 * nrm, (k, np.abs(np.sort(w[k]) - wr).max() / nrm)
        R = A @ U[k] - U[k] * w[k][None, :]
        assert np.abs(R).max() <= tol_vec * nrm * sc, (k, np.abs(R).max() / nrm)
        O = U[k].conj().T @ U[k] - np.eye(n)
        assert np.abs(O).max() <= tol_vec * sc, (k, np.abs(O).max())


@pytest.mark.parametrize("variant", ["auto", "valu", "lower", "fused"])
@pytest.mark.parametrize("n,b", [(1, 2), (2, 3), (3, 3), (31, 3), (33, 3), (64, 2), (96, 2), (97, 2), (128, 3),
                                 (129, 2), (160, 2), (161, 2), (200, 2), (224, 2), (225, 2), (256, 4), (300, 2),
                                 (384, 2)])
def test_mixed_precision(monkeypatch, variant, n, b):
    """HIP_DIAG_MIXED: FP32 tridiagonalisation (register-resident kernel for n <= 256 under
    "auto", with 1, 2 or 4 register tiles per wave and LDS tiles where 8 waves x slots fall
    short; its rank-2 update on the matrix cores, or on the VALU under "valu"
    (RLM_RESIDENT_VALU); above 256 panels, then the register-resident kernel on the trailing
    256 block) and back-transform, FP64 D&C and Rayleigh quotients."""
    import torch
    from rlm_amd import heevd_batched, HIP_DIAG_MIXED
    print(f"[test_mixed_precision] {variant} n={n} b={b}", flush=True)
    monkeypatch.delenv("RLM_RESIDENT_VALU", raising=False)
    if variant in ("auto", "valu"):
        monkeypatch.delenv("RLM_HETRD", raising=False)
        if variant == "valu":
            monkeypatch.setenv("RLM_RESIDENT_VALU", "1")
    else:
        monkeypatch.setenv("RLM_HETRD", variant)
    rng = np.random.default_rng(11 * n + b)
    H = _rand_herm(rng, b, n)
    w, U = heevd_batched(torch.from_numpy(H).cuda(), precision=HIP_DIAG_MIXED)
    _check_mixed(H, w.cpu().numpy(), U.cpu().numpy())


def test_mixed_precision_structured():
    """Mixed precision on zero, diagonal, banded-sparse and widely scaled matrices."""
    import torch
    from rlm_amd import heevd_batched, HIP_DIAG_MIXED
    print("[test_mixed_precision_structured]", flush=True)
    n = 256
    rng = np.random.default_rng(5)
    mats = [np.zeros((n, n), dtype=np.complex128),
            np.diag(np.linspace(-1, 1, n)).astype(np.complex128)]
    B = _rand_herm(rng, 1, n)[0]
    band = np.abs(np.subtract.outer(np.arange(n), np.arange(n)))
    B[(band > 33) & (band < n - 33)] = 0            # the circular band of the rlm Hamiltonian
    mats.append(B)
    mats.append(1e-30 * _rand_herm(rng, 1, n)[0])
    mats.append(1e30 * _rand_herm(rng, 1, n)[0])
    H = np.stack([(M + M.conj().T) / 2 for M in mats])
    w, U = heevd_batched(torch.from_numpy(H).cuda(), precision=HIP_DIAG_MIXED)
    _check_mixed(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("n", [64, 256])
def test_mixed_near_degenerate_clusters(n):
    """Mixed solver on H = Q diag(l) Q^H with a 5-fold and an 8-fold cluster of eigenvalues
    1e-7 apart (relative to an extent of 2) and a pair: the FP32 vectors mix inside each
    cluster, and a Rayleigh quotient would then err by |coupling|^2 / gap ~ 1e-7. The exact
    FP64 Rayleigh-Ritz of each run (up to 8 states in one chunk) must bring every
    eigenvalue within 1e-10 of the constructed spectrum. (Longer runs are split into
    chunks, whose couplings across the split remain: rlm_plan_pair_info counts them.)"""
    import torch
    from rlm_amd import heevd_batched, HIP_DIAG_MIXED
    print(f"[test_mixed_near_degenerate_clusters] n={n}", flush=True)
    rng = np.random.default_rng(n)
    lam = np.linspace(-1.0, 1.0, n)          # gaps 2/(n-1), far above 1e-4 of the extent
    for c0, m in ((n // 4, 5), (n // 2, 8), (3 * n // 4, 2)):
        lam[c0:c0 + m] = lam[c0] + 1e-7 * np.arange(m)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Q, _ = np.linalg.qr(X)
    H = (Q * lam[None, :]) @ Q.conj().T
    H = (H + H.conj().T) / 2
    w, U = heevd_batched(torch.from_numpy(H[None]).cuda(), precision=HIP_DIAG_MIXED)
    w = w.cpu().numpy()[0]
    wr = np.linalg.eigvalsh(H)
    assert np.abs(w - wr).max() <= 1e-10, np.abs(w - wr).max()
    U = U.cpu().numpy()[0]
    assert np.abs(H @ U - U * w[None, :]).max() <= 3e-5 * max(1, n / 64)
