"""CPU oracle for the rlm/hlm hot path (test infrastructure only).

A restatement of qiongz/Random-Landau-Model's double-precision CPU build
(``rlm_cpu_dp`` / ``hlm_cpu_dp``). The bit-sensitive scalar parts live in
``rlm_oracle.c`` (built by ``oracle/Makefile`` into ``liboracle_rlm.so``);
this module drives them and calls LAPACK ``zheevd`` through
``numpy.linalg.eigh`` exactly where the reference calls ``mkl_heevd``
(``diag_wrappers/mkl_diag.cpp:2-9``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_rlm.so")
_lib = None


def lib():
    """Load liboracle_rlm.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.orc_mt_words.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, u32p]
        L.orc_rlm_disorder.argtypes = [ctypes.c_uint32, ctypes.c_long, ctypes.c_long,
                                       ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_hlm_disorder.argtypes = [ctypes.c_uint32, ctypes.c_long, ctypes.c_long, dp]
        L.orc_prepare_coeff.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_potential.argtypes = [dp, dp, dp, ctypes.c_long, dp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_assemble.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   dp, dp, dp, dp]
        L.orc_chern_line.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_mt_init_by_array.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
        L.orc_mt_next.argtypes = [ctypes.c_void_p]
        L.orc_mt_next.restype = ctypes.c_uint32
        L.orc_mt_state_size.restype = ctypes.c_size_t
        L.orc_expf.argtypes = [ctypes.c_float]
        L.orc_expf.restype = ctypes.c_float
        for f in (L.orc_cosf, L.orc_sinf):
            f.argtypes = [ctypes.c_float]
            f.restype = ctypes.c_float
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# --------------------------------------------------------------------------
# RNG (MKL VSL MT19937 restated)
# --------------------------------------------------------------------------
def mt_words(seed: int, skip: int, count: int) -> np.ndarray:
    """Raw 32-bit words of vslNewStream(VSL_BRNG_MT19937, seed) after skip-ahead."""
    out = np.empty(count, dtype=np.uint32)
    lib().orc_mt_words(seed & 0xFFFFFFFF, skip, count, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


# --------------------------------------------------------------------------
# rlm sizes, tables, disorder, potential, assembly
# --------------------------------------------------------------------------
@dataclass(frozen=True)
class RlmSizes:
    n_phi: int
    L1: float
    L2: float
    n_imp: int
    off: int
    d: int


def rlm_sizes(n_phi: int, q: float = 1.0, imp: float = 16.0, off_override: int | None = None) -> RlmSizes:
    """rlm/main_rlm_cpu.cpp:29-34. ``off_override`` is the -d flag the reference documents."""
    L = math.sqrt(n_phi * q)
    n_imp = int(imp * n_phi / q)
    off = int(math.sqrt(float(n_imp)) / 2 + 1) if off_override is None else int(off_override)
    return RlmSizes(n_phi, L, L, n_imp, off, 2 * off + 1)


def rlm_tables(sz: RlmSizes, n_mesh: int):
    """rlm/matrix_coefficients.cpp:3-25 prepare_coeff -> (coeff_mn[d,d], coeff_m_theta[g+1,d], coeff_jm[N,d])."""
    d = sz.d
    cmn = np.zeros((d, d), dtype=np.complex128)
    cmt = np.zeros((n_mesh + 1, d), dtype=np.complex128)
    cjm = np.zeros((sz.n_phi, d), dtype=np.complex128)
    lib().orc_prepare_coeff(sz.n_phi, sz.off, d, n_mesh, sz.L1, sz.L2,
                            _dp(cmn.view(np.float64)), _dp(cmt.view(np.float64)), _dp(cjm.view(np.float64)))
    return cmn, cmt, cjm


# --------------------------------------------------------------------------
# The reference's single-precision GPU builds (rlm_gpu_*: dtype float, writing
# energy_Chern_No.dat). Their conventions, restated; arithmetic after the float
# inputs in FP64.
# --------------------------------------------------------------------------
def rlm_sizes_f32(n_phi: int, q: float = 1.0, imp: float = 16.0) -> RlmSizes:
    """rlm/main_rlm_gpu.cu:29-34: float L1 = L2 = sqrt(n_phi*q), float impurity count."""
    f = np.float32
    L = float(np.sqrt(f(n_phi) * f(q)))
    n_imp = int(f(imp) * f(n_phi) / f(q))
    off = int(math.sqrt(float(n_imp)) / 2 + 1)
    return RlmSizes(n_phi, L, L, n_imp, off, 2 * off + 1)


def rlm_tables_f32(sz: RlmSizes, n_mesh: int):
    """rlm/matrix_coefficients.cpp:3-25 compiled with dtype float -> complex tables
    (float values as complex128)."""
    f = np.float32
    L = lib()
    N, o, d = sz.n_phi, sz.off, sz.d
    L1, L2 = f(sz.L1), f(sz.L2)
    cmn = np.zeros((d, d), dtype=np.complex128)
    for m in range(d):
        for n in range(d):
            ph = math.pi * (m - o) * (n - o) / N
            q1 = f((m - o) * (m - o)) * L2 / L1
            q2 = f((n - o) * (n - o)) * L1 / L2
            arg = -math.pi * float(f(q1 + q2)) / 2.0 / N
            e = f(L.orc_expf(float(f(arg))))
            r = f(1.0) / f(N)
            cmn[m, n] = complex(float((f(math.cos(ph)) * e) * r), float((f(-math.sin(ph)) * e) * r))
    cmt = np.zeros((n_mesh + 1, d), dtype=np.complex128)
    for t in range(n_mesh + 1):
        dth = f(t * (math.pi * 2.0) / n_mesh)
        for m in range(d):
            a = f(f(m - o) * dth) / f(N)
            cmt[t, m] = complex(L.orc_cosf(float(a)), L.orc_sinf(float(a)))
    cjm = np.zeros((N, d), dtype=np.complex128)
    for j in range(N):
        for m in range(d):
            a = (math.pi * 2.0) * (m - o) * (j + 1) / N
            cjm[j, m] = complex(float(f(math.cos(a))), float(f(-math.sin(a))))
    return cmn, cmt, cjm


def rlm_disorder_f32(sz: RlmSizes, seed: int, sample: int):
    """rlm/disorder_potential.cpp:3-21 with dtype float: vsRngUniform gives
    L * float(word) * 2^-32 in float arithmetic (pinned against MKL)."""
    f = np.float32
    w = mt_words(seed, 3 * sz.n_imp * sample, 3 * sz.n_imp)
    n = sz.n_imp
    x = (f(sz.L1) * (w[:n].astype(np.float32) * f(2.0 ** -32))).astype(np.float64)
    y = (f(sz.L2) * (w[n:2 * n].astype(np.float32) * f(2.0 ** -32))).astype(np.float64)
    s = np.where(w[2 * n:] % 2 == 0, 1.0, -1.0)
    return x, y, s


def rlm_sample_gpu_conv(n_phi: int, q: float, imp: float, n_mesh: int, seed: int, sample: int,
                        threads: int = 1) -> SampleResult:
    """One sample of the FP32 GPU builds' pipeline (rlm/main_rlm_gpu.cu:138-300) with FP64
    arithmetic after the float inputs: V_mn = coeff_mn sum_i s_i e^{+i phase}
    (rlm/potential_coeff.cu:14-15) and h_jk = sum_m coeff_jm[j][m] c2[m] sum_n V[m][n] conj(c1[n])
    (rlm/hamiltonian_rlm_gpu.cu:23,36): the CPU formulas of rlm_potential / rlm_assemble
    evaluated at negated positions and with conj(coeff_m_theta). Energies, Chern sums and
    the +1/N offset as the CPU build (main_rlm_gpu.cu:230-297 has the same loops)."""
    sz = rlm_sizes_f32(n_phi, q, imp)
    cmn, cmt, cjm = rlm_tables_f32(sz, n_mesh)
    x, y, s = rlm_disorder_f32(sz, seed, sample)
    V = rlm_potential(sz, -x, -y, s, cmn)
    return rlm_sample(n_phi, q, imp, n_mesh, seed, sample, threads=threads,
                      tables=(cmn, np.conj(cmt), cjm), sizes=sz, potential=V)


def rlm_disorder(sz: RlmSizes, seed: int, sample: int):
    """rlm/disorder_potential.cpp:3-21 -> (x, y, s)."""
    x = np.empty(sz.n_imp)
    y = np.empty(sz.n_imp)
    s = np.empty(sz.n_imp)
    lib().orc_rlm_disorder(seed & 0xFFFFFFFF, sample, sz.n_imp, sz.L1, sz.L2, _dp(x), _dp(y), _dp(s))
    return x, y, s


def rlm_potential(sz: RlmSizes, x, y, s, cmn) -> np.ndarray:
    """rlm/matrix_coefficients.cpp:27-44 get_potential_coeff -> V[d,d]."""
    V = np.zeros((sz.d, sz.d), dtype=np.complex128)
    lib().orc_potential(_dp(np.ascontiguousarray(x)), _dp(np.ascontiguousarray(y)), _dp(np.ascontiguousarray(s)),
                        sz.n_imp, _dp(np.ascontiguousarray(cmn).view(np.float64)), sz.d, sz.off,
                        _dp(V.view(np.float64)))
    return V


def rlm_assemble_buffer(sz: RlmSizes, V, cmt, cjm, theta_1: int, theta_2: int) -> np.ndarray:
    """rlm/hamiltonian_rlm_cpu.cpp:14-48: the n_phi*n_phi buffer as the reference writes it."""
    wf = np.zeros(sz.n_phi * sz.n_phi, dtype=np.complex128)
    lib().orc_assemble(sz.n_phi, sz.d, sz.off, theta_1, theta_2, _dp(np.ascontiguousarray(V).view(np.float64)),
                       _dp(np.ascontiguousarray(cmt).view(np.float64)),
                       _dp(np.ascontiguousarray(cjm).view(np.float64)), _dp(wf.view(np.float64)))
    return wf


def lapack_view(buf: np.ndarray, n: int) -> np.ndarray:
    """The matrix LAPACK sees when handed ``buf`` column-major with lda=n: A[r,c] = buf[c*n+r]."""
    return buf.reshape(n, n).T


def rlm_hamiltonian(sz: RlmSizes, V, cmt, cjm, theta_1: int, theta_2: int) -> np.ndarray:
    """The Hermitian matrix A (A[r,c]) that zheevd('V','U') diagonalises at (theta_1, theta_2)."""
    return np.ascontiguousarray(lapack_view(rlm_assemble_buffer(sz, V, cmt, cjm, theta_1, theta_2), sz.n_phi))


def heevd(A: np.ndarray):
    """diag_wrappers/mkl_diag.cpp:2-9 mkl_heevd: LAPACK zheevd, jobz='V', uplo='U'.

    Returns (w ascending, Z) with Z[:, n] the eigenvector of state n.
    """
    return np.linalg.eigh(A, UPLO="U")


def chern_line(ring: np.ndarray, n_phi: int, dim_vec: int, n_mesh: int, theta_1: int) -> np.ndarray:
    """diag_wrappers/chern.cpp:37-75 over theta_2 = 0..n_mesh-1.

    ring: complex array shaped [2, n_mesh+1, n_phi, dim_vec] (slot, theta_2, state, component).
    Returns chern_theta[n_phi * n_mesh] with index m + n*n_mesh.
    """
    ring = np.ascontiguousarray(ring, dtype=np.complex128)
    out = np.zeros(n_phi * n_mesh)
    lib().orc_chern_line(_dp(ring.view(np.float64).reshape(-1)), n_phi, dim_vec, n_mesh, theta_1, _dp(out))
    return out


@dataclass
class SampleResult:
    energy: np.ndarray          # mesh-averaged eigenenergies [n_phi]
    chern: np.ndarray           # per-state Chern values (float) [n_phi]
    energy_theta: np.ndarray    # [g+1, g+1, n_phi] eigenvalues per twist point (rlm) / [g+1, g, n_phi] (hlm)


def _solve_line_rlm(sz, V, cmt, cjm, theta_1, n_mesh, pool):
    def one(t2):
        A = rlm_hamiltonian(sz, V, cmt, cjm, theta_1, t2)
        return heevd(A)
    if pool is None:
        res = [one(t2) for t2 in range(n_mesh + 1)]
    else:
        res = list(pool.map(one, range(n_mesh + 1)))
    w = np.stack([r[0] for r in res])                       # [g+1, N]
    vecs = np.stack([r[1].T for r in res])                  # [g+1, state, component]
    return w, vecs


def rlm_sample(n_phi: int, q: float, imp: float, n_mesh: int, seed: int, sample: int,
               threads: int = 1, off_override: int | None = None, tables=None, sizes=None,
               potential=None) -> SampleResult:
    """rlm/main_rlm_cpu.cpp:109-243 for one sample (same accumulation order)."""
    sz = sizes if sizes is not None else rlm_sizes(n_phi, q, imp, off_override)
    cmn, cmt, cjm = tables if tables is not None else rlm_tables(sz, n_mesh)
    if potential is None:
        x, y, s = rlm_disorder(sz, seed, sample)
        V = rlm_potential(sz, x, y, s, cmn)
    else:
        V = potential
    pool = ThreadPoolExecutor(threads) if threads > 1 else None
    try:
        N = n_phi
        energy = np.zeros(N)
        chern = np.zeros(N)
        ring = np.zeros((2, n_mesh + 1, N, N), dtype=np.complex128)
        e_theta_all = np.zeros((n_mesh + 1, n_mesh + 1, N))
        w, vecs = _solve_line_rlm(sz, V, cmt, cjm, 0, n_mesh, pool)
        ring[0] = vecs
        e_theta_all[0] = w
        for i in range(N):                                   # :155-158
            for j in range(n_mesh + 1):
                energy[i] += w[j, i]
        for theta_1 in range(1, n_mesh + 1):                 # :161
            w, vecs = _solve_line_rlm(sz, V, cmt, cjm, theta_1, n_mesh, pool)
            ring[theta_1 % 2] = vecs
            e_theta_all[theta_1] = w
            for i in range(N):                               # :187-190
                for j in range(n_mesh + 1):
                    energy[i] += w[j, i]
            ct = chern_line(ring, N, N, n_mesh, theta_1)     # :195-208
            for k in range(N):                               # :230-233
                for i in range(n_mesh):
                    chern[k] += ct[k * n_mesh + i]
        for k in range(N):                                   # :236-239
            energy[k] /= (n_mesh + 1) * (n_mesh + 1)
            chern[k] += 1.0 / n_phi
        return SampleResult(energy, chern, e_theta_all)
    finally:
        if pool is not None:
            pool.shutdown()


# --------------------------------------------------------------------------
# hlm
# --------------------------------------------------------------------------
def hlm_disorder(seed: int, n_sites: int, n_samples: int) -> np.ndarray:
    """hlm/main_hlm_cpu.cpp:66-67,79 -> [n_samples, n_sites] values in [-1, 1)."""
    out = np.empty(n_sites * n_samples)
    lib().orc_hlm_disorder(seed & 0xFFFFFFFF, n_sites, n_samples, _dp(out))
    return out.reshape(n_samples, n_sites)


def hlm_bloch_matrix(Q: int, kx_shifted: float, ky_shifted: float, lx: int, ly: int) -> np.ndarray:
    """hlm/hamiltonian_hlm_cpu.cpp:26-32: the QxQ Bloch matrix as LAPACK sees it (column-major writes)."""
    PI2 = 2.0 * math.pi
    pi2_inv_lx = PI2 / lx
    pi2_inv_ly = PI2 / ly
    inv_Q = 1.0 / Q
    h = np.zeros(Q * Q, dtype=np.complex128)
    for p in range(Q):
        h[p * Q + p] = -2.0 * math.cos(ky_shifted * pi2_inv_ly + PI2 * p * inv_Q)
        if p + 1 < Q:
            h[p * Q + p + 1] = -1.0
            h[(p + 1) * Q + p] = -1.0
    ph = kx_shifted * Q * pi2_inv_lx
    h[Q * (Q - 1)] = -complex(math.cos(ph), math.sin(ph))
    h[Q - 1] = np.conj(h[Q * (Q - 1)])
    return lapack_view(h, Q)


def hlm_basis(lx: int, ly: int, Q: int, band: int, n_mesh: int, theta_1: int, theta_2: int) -> np.ndarray:
    """hlm/hamiltonian_hlm_cpu.cpp:2-56 get_wfs_ktheta -> Phi[n_sites, n_phi] (column kx*ly+ky)."""
    n_sites = lx * ly
    n_phi = n_sites // Q
    PI2 = 2.0 * math.pi
    inv_n_mesh = 1.0 / n_mesh
    pi2_inv_lx = PI2 / lx
    pi2_inv_ly = PI2 / ly
    Phi = np.zeros((n_sites, n_phi), dtype=np.complex128)
    for kx in range(lx // Q):
        kxs = kx + theta_1 * inv_n_mesh
        for ky in range(ly):
            kys = ky + theta_2 * inv_n_mesh
            col = kx * ly + ky
            _, Z = heevd(hlm_bloch_matrix(Q, kxs, kys, lx, ly))
            vec = Z[:, band]
            w = np.zeros(n_sites, dtype=np.complex128)
            for j in range(ly):
                for m in range(0, lx, Q):
                    phase = kxs * pi2_inv_lx * m + kys * pi2_inv_ly * j
                    c = complex(math.cos(phase), -math.sin(phase))
                    for qq in range(Q):
                        w[j * lx + m + qq] = c * vec[qq]
            norm = float(np.sum(w.real * w.real + w.imag * w.imag))
            w *= 1.0 / math.sqrt(norm)
            Phi[:, col] = w
    return Phi


def hlm_solve_twist(lx, ly, Q, band, n_mesh, potential, theta_1, theta_2):
    """hlm/hamiltonian_hlm_cpu.cpp:149-184 for one twist point -> (w[n_phi], wfs[n_phi, n_sites])."""
    Phi = hlm_basis(lx, ly, Q, band, n_mesh, theta_1, theta_2)
    W = potential[:, None] * Phi
    Htr = Phi.conj().T @ W
    w, U = heevd(Htr)
    wfs = Phi @ U                                  # [n_sites, n_phi]
    return w, np.ascontiguousarray(wfs.T)          # states contiguous


def hlm_sample(lx: int, ly: int, Q: int, band: int, n_mesh: int, potential: np.ndarray,
               threads: int = 1) -> SampleResult:
    """hlm/main_hlm_cpu.cpp:75-196 for one sample's potential (same accumulation order)."""
    n_sites = lx * ly
    N = n_sites // Q
    pool = ThreadPoolExecutor(threads) if threads > 1 else None

    def line(theta_1):
        f = lambda t2: hlm_solve_twist(lx, ly, Q, band, n_mesh, potential, theta_1, t2)
        res = list(pool.map(f, range(n_mesh))) if pool else [f(t2) for t2 in range(n_mesh)]
        return np.stack([r[0] for r in res]), np.stack([r[1] for r in res])

    try:
        energy = np.zeros(N)
        chern = np.zeros(N)
        ring = np.zeros((2, n_mesh + 1, N, n_sites), dtype=np.complex128)
        e_all = np.zeros((n_mesh + 1, n_mesh, N))
        w, vecs = line(0)
        e_all[0] = w
        ring[0, :n_mesh] = vecs
        for i in range(N):                                   # :111-115
            for j in range(n_mesh):
                energy[i] += w[j, i]
            energy[i] += w[0, i]
        ring[0, n_mesh] = ring[0, 0]                         # :117-118
        for theta_1 in range(1, n_mesh + 1):
            w, vecs = line(theta_1)
            e_all[theta_1] = w
            slot = theta_1 % 2
            ring[slot, :n_mesh] = vecs
            for i in range(N):                               # :143-147
                for j in range(n_mesh):
                    energy[i] += w[j, i]
                energy[i] += w[0, i]
            ring[slot, n_mesh] = ring[slot, 0]               # :149-150
            ct = chern_line(ring, N, n_sites, n_mesh, theta_1)
            for k in range(N):                               # :184-187
                for i in range(n_mesh):
                    chern[k] += -ct[k * n_mesh + i]
        for k in range(N):                                   # :190-191
            energy[k] /= (n_mesh + 1) * (n_mesh + 1)
        return SampleResult(energy, chern, e_all)
    finally:
        if pool is not None:
            pool.shutdown()


# --------------------------------------------------------------------------
# energy_Chern_No.txt formatting (rlm/main_rlm_cpu.cpp:242, hlm/main_hlm_cpu.cpp:195)
# --------------------------------------------------------------------------
def _g8(v: float) -> str:
    """C++ ostream default floatfield with setprecision(8) (printf %.8g)."""
    return "%.8g" % v


def format_rows(sample: int, energy, chern, left: bool) -> str:
    lines = []
    for e, c in zip(energy, chern):
        if left:
            lines.append(f"{str(sample):<8} {_g8(e):<15} {_g8(c)}")
        else:
            lines.append(f"{str(sample):>8} {_g8(e):>15} {_g8(c)}")
    return "\n".join(lines) + "\n"
