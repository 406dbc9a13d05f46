"""Pins the oracle's FHS Chern reduction (diag_wrappers/chern.cpp:37-75) and the hlm
Bloch basis (hlm/hamiltonian_hlm_cpu.cpp:2-56) against textbook results:

* the Qi-Wu-Zhang two-band model, whose band Chern numbers follow from the
  continuum projector formula C = (1/2pi) int i Tr(P [d1 P, d2 P]) computed here by
  finite differences (independent of the lattice link formula);
* gauge invariance of the plaquette phases;
* the real-space twin get_wfs_rspace (hlm/hamiltonian_hlm_cpu.cpp:113-147), which the
  reference keeps to "check if the k-space transformated wfs is correct" (:168-170).
"""
import math

import numpy as np
import pytest


def _qwz(kx, ky, u):
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]], dtype=complex)
    sz = np.array([[1, 0], [0, -1]], dtype=complex)
    return math.sin(kx) * sx + math.sin(ky) * sy + (u + math.cos(kx) + math.cos(ky)) * sz


def _chern_oracle(oracle, u, g, phases=None):
    """Drive oracle.chern_line exactly as rlm/main_rlm_cpu.cpp:161-233 does (no +1/N)."""
    vecs = np.zeros((g + 1, g + 1, 2, 2), dtype=complex)     # [t1][t2][state][comp]
    for t1 in range(g + 1):
        for t2 in range(g + 1):
            _, Z = np.linalg.eigh(_qwz(2 * np.pi * t1 / g, 2 * np.pi * t2 / g, u))
            vecs[t1, t2] = Z.T
    if phases is not None:
        vecs = vecs * phases[..., None]
    ring = np.zeros((2, g + 1, 2, 2), dtype=complex)
    ring[0] = vecs[0]
    chern = np.zeros(2)
    for t1 in range(1, g + 1):
        ring[t1 % 2] = vecs[t1]
        ct = oracle.chern_line(ring, 2, 2, g, t1)
        for k in range(2):
            for i in range(g):
                chern[k] += ct[k * g + i]
    return chern


def _chern_projector(u, M=160):
    """(1/2pi) sum i Tr(P [dx P, dy P]) dk^2 for the lower band, central differences."""
    h = 2 * np.pi / M
    ks = np.arange(M) * h

    def P(kx, ky):
        _, Z = np.linalg.eigh(_qwz(kx, ky, u))
        v = Z[:, 0]
        return np.outer(v, v.conj())
    tot = 0.0
    for kx in ks:
        for ky in ks:
            p = P(kx, ky)
            dx = (P(kx + h / 2, ky) - P(kx - h / 2, ky)) / h
            dy = (P(kx, ky + h / 2) - P(kx, ky - h / 2)) / h
            tot += (1j * np.trace(p @ (dx @ dy - dy @ dx))).real
    return tot * h * h / (2 * np.pi)


@pytest.mark.parametrize("u", [1.0, -1.0, 3.0])
def test_qwz_matches_projector_formula(oracle, u):
    c = _chern_oracle(oracle, u, 24)
    cp = _chern_projector(u)
    assert abs(cp - round(cp)) < 2e-2
    assert np.allclose(c, np.rint(c), atol=1e-10)
    assert round(c[0]) == round(cp)
    assert round(c[0]) + round(c[1]) == 0
    if abs(u) < 2:
        assert abs(round(c[0])) == 1


def test_gauge_invariance(oracle):
    g = 12
    rng = np.random.default_rng(0)
    ph = np.exp(2j * np.pi * rng.random((g + 1, g + 1, 2)))
    c0 = _chern_oracle(oracle, 1.0, g)
    c1 = _chern_oracle(oracle, 1.0, g, ph)
    assert np.abs(c0 - c1).max() < 1e-12


def _rspace_hamiltonian(lx, ly, Q, n_mesh, t1, t2):
    """hlm/hamiltonian_hlm_cpu.cpp:113-147 get_wfs_rspace, the matrix it hands mkl_heevd."""
    n = lx * ly
    buf = np.zeros(n * n, dtype=complex)
    P2 = 2 * math.pi
    for i in range(lx):
        for j in range(ly):
            ir = i + 1 - lx if i + 1 >= lx else i + 1
            il = i - 1 + lx if i - 1 < 0 else i - 1
            buf[(i * ly + j) * n + ir * ly + j] = -1
            buf[(ir * ly + j) * n + i * ly + j] = -1
            buf[(i * ly + j) * n + il * ly + j] = -1
            buf[(il * ly + j) * n + i * ly + j] = -1
            jr = j + 1 - ly if j + 1 >= ly else j + 1
            jl = j - 1 + ly if j - 1 < 0 else j - 1
            c, s = math.cos(P2 * i / Q), math.sin(P2 * i / Q)
            buf[(i * ly + j) * n + i * ly + jr] = -complex(c, s)
            buf[(i * ly + jr) * n + i * ly + j] = -complex(c, -s)
            buf[(i * ly + j) * n + i * ly + jl] = -complex(c, -s)
            buf[(i * ly + jl) * n + i * ly + j] = -complex(c, s)
    for i in range(lx):
        a = P2 * i / Q + t2 * P2 / n_mesh
        buf[(i * ly + ly - 1) * n + i * ly] = -complex(math.cos(a), math.sin(a))
        buf[(i * ly) * n + i * ly + ly - 1] = -complex(math.cos(a), -math.sin(a))
    for j in range(ly):
        a = t1 * P2 / n_mesh
        buf[((lx - 1) * ly + j) * n + j] = -complex(math.cos(a), math.sin(a))
        buf[j * n + (lx - 1) * ly + j] = -complex(math.cos(a), -math.sin(a))
    return buf.reshape(n, n).T


@pytest.mark.parametrize("lx,ly,Q,t1,t2", [(6, 6, 3, 0, 0), (6, 6, 3, 2, 5), (8, 4, 4, 3, 1), (9, 6, 3, 7, 7)])
def test_bloch_basis_vs_real_space(oracle, lx, ly, Q, t1, t2):
    """The projected disorder spectrum from the k-space basis equals the one from the
    lowest n_phi real-space eigenvectors (the twist gauges differ by a diagonal unitary,
    which commutes with the on-site potential)."""
    g = 8
    n_sites = lx * ly
    n_phi = n_sites // Q
    rng = np.random.default_rng(lx * 100 + t1 * 10 + t2)
    V = rng.uniform(-1, 1, n_sites)
    Phi = oracle.hlm_basis(lx, ly, Q, 0, g, t1, t2)                  # sites r = y*lx + x
    assert np.allclose(Phi.conj().T @ Phi, np.eye(n_phi), atol=1e-12)
    e_k = np.linalg.eigvalsh(Phi.conj().T @ (V[:, None] * Phi))
    Hr = _rspace_hamiltonian(lx, ly, Q, g, t1, t2)                    # sites i*ly + j (x*ly + y)
    w, Z = np.linalg.eigh(Hr)
    assert w[n_phi] - w[n_phi - 1] > 1e-3                            # lowest subband is isolated
    Zb = Z[:, :n_phi]
    perm = np.array([(r % lx) * ly + r // lx for r in range(n_sites)])  # k-space site -> real-space site
    Vr = np.empty(n_sites)
    Vr[perm] = V
    e_r = np.linalg.eigvalsh(Zb.conj().T @ (Vr[:, None] * Zb))
    assert np.abs(e_k - e_r).max() < 1e-10
