"""Pins the oracle's rlm model arithmetic (prepare_coeff, get_potential_coeff, the
assembly loop) against independent restatements: libm called directly, brute-force
DFTs, the reference's own brute-force Kronecker-delta loop
(rlm/hamiltonian_rlm_cpu.cpp:40-43, commented "old version") and closed-form
invariants (trace, twist periodicity)."""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

libm = ctypes.CDLL(ctypes.util.find_library("m"))
libm.expf.argtypes = [ctypes.c_float]
libm.expf.restype = ctypes.c_float


def _ulp_close(a: complex, b: complex, ulps: int = 2) -> bool:
    """Components agree to `ulps` units in the last place (glibc's cos/sin and
    sincos -- which compilers merge them into -- may differ by an ulp)."""
    for u, v in ((a.real, b.real), (a.imag, b.imag)):
        if abs(u - v) > ulps * np.spacing(max(abs(u), abs(v), 1e-300)):
            return False
    return True


@pytest.mark.parametrize("n_phi,g", [(16, 8), (64, 16), (192, 4), (256, 32)])
def test_prepare_coeff_closed_form(oracle, n_phi, g):
    sz = oracle.rlm_sizes(n_phi)
    cmn, cmt, cjm = oracle.rlm_tables(sz, g)
    o, d, N = sz.off, sz.d, n_phi
    L1 = L2 = sz.L1
    f32_inv = float(np.float32(1.0) / np.float32(N))
    used_f32 = False
    for m in range(d):
        for n in range(d):
            ph = math.pi * (m - o) * (n - o) / N
            arg = -math.pi * ((m - o) * (m - o) * L2 / L1 + (n - o) * (n - o) * L1 / L2) / 2.0 / N
            e = float(libm.expf(float(np.float32(arg))))
            want = complex(math.cos(ph) * e * f32_inv, -math.sin(ph) * e * f32_inv)
            assert _ulp_close(cmn[m, n], want), (m, n)
            # the exp factor itself is exactly float32 expf(float32(arg)) * float32(1/N)
            assert abs(abs(cmn[m, n]) - e * f32_inv) <= 4e-16 * e * f32_inv
            dbl = complex(math.cos(ph), -math.sin(ph)) * math.exp(arg) / N
            used_f32 |= (cmn[m, n] != dbl)
    assert used_f32, "the single-precision expf / 1.0f/n_phi island must show in the table"
    for t in range(g + 1):
        dth = t * (math.pi * 2.0) / g
        for m in range(d):
            a = (m - o) * dth / N
            assert _ulp_close(cmt[t, m], complex(math.cos(a), math.sin(a)))
    for j in range(N):
        for m in range(d):
            a = (math.pi * 2.0) * (m - o) * (j + 1) / N
            assert _ulp_close(cjm[j, m], complex(math.cos(a), -math.sin(a)))


@pytest.mark.parametrize("n_phi,seed,sample", [(16, 1, 0), (64, 11, 3), (256, 7, 1)])
def test_potential_brute_force(oracle, n_phi, seed, sample):
    sz = oracle.rlm_sizes(n_phi)
    cmn, _, _ = oracle.rlm_tables(sz, 4)
    x, y, s = oracle.rlm_disorder(sz, seed, sample)
    V = oracle.rlm_potential(sz, x, y, s, cmn)
    k = 2 * np.pi * (np.arange(sz.d) - sz.off)
    ex = np.exp(-1j * np.outer(k, x))          # [m, i]
    ey = np.exp(-1j * np.outer(k, y))          # [n, i]
    Vb = (ex * s[None, :]) @ ey.T * cmn
    assert np.abs(V - Vb).max() <= 1e-11 * np.abs(V).max()
    # V[o][o] = coeff_mn[o][o] * sum(s) exactly (zero phase)
    assert V[sz.off, sz.off] == cmn[sz.off, sz.off] * s.sum()


def _kronecker_assembly(sz, V, cmt, cjm, t1, t2):
    """rlm/hamiltonian_rlm_cpu.cpp:40-43 'old version': sum over every n with |j-k+n-o| % N == 0."""
    N, d, o = sz.n_phi, sz.d, sz.off
    h = np.zeros((N, N), dtype=np.complex128)
    for j in range(N):
        ct = np.conj(cmt[t2]) * cjm[j]                         # coeff_theta_jm over m
        for k in range(j, N):
            delta = j - k - o
            ns = [n for n in range(d) if abs(delta + n) % N == 0]
            S = np.zeros(d, dtype=np.complex128)
            for n in ns:
                S = S + V[:, n] * cmt[t1, n]
            h[j, k] = np.sum(S * ct)
            h[k, j] = np.conj(h[j, k])
    # the reference writes conj(h) last at [k*N+j] for k=j; LAPACK sees A = h^T
    return h.T.copy()


@pytest.mark.parametrize("n_phi,g,seed,t1,t2", [(16, 8, 1, 3, 5), (8, 4, 2, 1, 4), (40, 6, 4, 6, 0)])
def test_assembly_vs_kronecker_loop(oracle, n_phi, g, seed, t1, t2):
    sz = oracle.rlm_sizes(n_phi)
    cmn, cmt, cjm = oracle.rlm_tables(sz, g)
    x, y, s = oracle.rlm_disorder(sz, seed, 0)
    V = oracle.rlm_potential(sz, x, y, s, cmn)
    A = oracle.rlm_hamiltonian(sz, V, cmt, cjm, t1, t2)
    B = _kronecker_assembly(sz, V, cmt, cjm, t1, t2)
    assert np.abs(A - B).max() <= 1e-13 * np.abs(A).max()
    # Hermitian off the diagonal bit for bit; the diagonal keeps a rounding-level
    # imaginary part, which zheevd ignores
    off = ~np.eye(sz.n_phi, dtype=bool)
    assert np.array_equal(A[off], A.conj().T[off])
    assert np.abs(np.diag(A).imag).max() <= 1e-14 * np.abs(A).max()


@pytest.mark.parametrize("n_phi,g,seed", [(16, 6, 1), (64, 4, 11), (96, 3, 5)])
def test_trace_and_twist_periodicity(oracle, n_phi, g, seed):
    """Tr H = N*fl32(1/N)*sum(s) for N > off_head; spectrum periodic in both twists."""
    sz = oracle.rlm_sizes(n_phi)
    cmn, cmt, cjm = oracle.rlm_tables(sz, g)
    x, y, s = oracle.rlm_disorder(sz, seed, 1)
    V = oracle.rlm_potential(sz, x, y, s, cmn)
    tr = n_phi * float(np.float32(1.0) / np.float32(n_phi)) * s.sum()
    spec = {}
    for t1 in (0, g):
        for t2 in (0, 1, g):
            A = oracle.rlm_hamiltonian(sz, V, cmt, cjm, t1, t2)
            assert abs(np.trace(A).real - tr) <= 1e-12 * max(1.0, abs(tr)) * n_phi
            spec[(t1, t2)] = np.linalg.eigvalsh(A)
    sc = np.abs(spec[(0, 0)]).max()
    assert np.abs(spec[(g, 0)] - spec[(0, 0)]).max() <= 1e-12 * sc
    assert np.abs(spec[(0, g)] - spec[(0, 0)]).max() <= 1e-12 * sc
    assert np.abs(spec[(0, 1)] - spec[(0, 0)]).max() > 1e-6 * sc   # the twist does act


def test_format_rows_matches_readme(oracle):
    # README.md rlm example rows: "0        -5.7376113      -1.2490009e-16"
    out = oracle.format_rows(0, [-5.7376113], [-1.2490009e-16], left=True)
    assert out == "0        -5.7376113      -1.2490009e-16\n"
    out = oracle.format_rows(0, [0.2389036], [4.3042826e-17], left=False)
    # README.md hlm example row "       0       0.2389036 4.3042826e-17"
    assert out == "       0       0.2389036 4.3042826e-17\n"
