"""GPU parity of the multi-workgroup eigensolver for 512 < n <= 4096 (eig_large.hip) and of
the rlm pipeline at those sizes, up to the n_phi = 3000 configuration (-n 3000 -q 1 -i 16,
g = 4, seed 3).

Bars: FP64 eigenvalues, residuals and orthogonality within 1e-12 * ||A|| * n/64 of LAPACK
zheevd (numpy.linalg.eigh, the routine of diag_wrappers/mkl_diag.cpp:8); mixed precision
(FP32 reduction and back-transform, FP64 divide & conquer and Rayleigh quotients)
eigenvalues within 1e-9 * ||A||, vectors within 3e-5 * n/64; rlm pipelines: Chern integers
identical to the oracle's, energies within 1e-10 relative.
"""
import numpy as np
import pytest

from test_gpu_eig import _check, _check_mixed, _rand_herm
from test_gpu_rlm import _cmp_sample

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,b", [(513, 2), (600, 3), (1000, 1)])
def test_large_random_hermitian_fp64(n, b):
    import torch
    from rlm_amd import heevd_batched
    print(f"[test_large_random_hermitian_fp64] n={n} b={b}", flush=True)
    rng = np.random.default_rng(300 + n)
    H = _rand_herm(rng, b, n)
    w, U = heevd_batched(torch.from_numpy(H).cuda())
    _check(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("n,b", [(513, 2), (777, 2)])
def test_large_random_hermitian_mixed(n, b):
    import torch
    from rlm_amd import heevd_batched, HIP_DIAG_MIXED
    print(f"[test_large_random_hermitian_mixed] n={n} b={b}", flush=True)
    rng = np.random.default_rng(500 + n)
    H = _rand_herm(rng, b, n)
    w, U = heevd_batched(torch.from_numpy(H).cuda(), precision=HIP_DIAG_MIXED)
    _check_mixed(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("n", [600, 1030])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_large_structured(precision, n):
    """Zero, identity (fully degenerate: every merge deflates), diagonal, circular-band
    (the rlm Hamiltonian's sparsity at n = 600, off_head = 49) and 1e+-30-scaled matrices;
    at n = 1030 the top merge's eigenvector product runs on several workgroups per matrix."""
    import torch
    from rlm_amd import heevd_batched, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print(f"[test_large_structured] {precision} n={n}", flush=True)
    rng = np.random.default_rng(17)
    mats = [np.zeros((n, n), dtype=np.complex128), np.eye(n, dtype=np.complex128),
            np.diag(np.linspace(-1, 1, n)).astype(np.complex128)]
    B = _rand_herm(rng, 1, n)[0]
    band = np.abs(np.subtract.outer(np.arange(n), np.arange(n)))
    B[(band > 49) & (band < n - 49)] = 0
    mats.append(B)
    mats.append(1e-30 * _rand_herm(rng, 1, n)[0])
    mats.append(1e30 * _rand_herm(rng, 1, n)[0])
    H = np.stack([(M + M.conj().T) / 2 for M in mats])
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    w, U = heevd_batched(torch.from_numpy(H).cuda(), precision=prec)
    (_check_mixed if precision == "mixed" else _check)(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rlm_large_vs_live_oracle(oracle, precision):
    """rlm at n_phi = 600 (-q 1 -i 16, off_head = 49), g = 2, seed 3, sample 0 through the
    multi-workgroup solver, against the CPU oracle."""
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print(f"[test_rlm_large_vs_live_oracle] n=600 g=2 {precision}", flush=True)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    plan = RlmPlan(600, 1.0, 16.0, 2, 3, samples_per_batch=1, precision=prec)
    e, c, m = plan.run(0, 1)
    plan.close()
    r = oracle.rlm_sample(600, 1.0, 16.0, 2, 3, 0, threads=8)
    assert m.min() > 1e-3, m.min()
    _cmp_sample(e[0], c[0], r.energy, r.chern, (600, 2, precision))


def _run_n3000(precision):
    from rlm_amd import RlmPlan, HIP_DIAG_MIXED
    plan = RlmPlan(3000, 1.0, 16.0, 4, 3, samples_per_batch=4, precision=precision)
    e, c, m = plan.run(0, 4)
    res = plan.max_residual()
    if precision == HIP_DIAG_MIXED:
        refines, cut = plan.refine_info()
        print(f"[n3000] mixed plan refines {refines}, chunks cut or unfactorised {cut}", flush=True)
        assert refines
    plan.close()
    return e, c, m, res


def test_rlm_n3000_g4_fp64():
    """The n_phi = 3000 configuration (-n 3000 -q 1 -i 16 -g 4, seed 3, 4 samples; 3000 x 3000
    complex matrices, off_head = 110) on the FP64 plan: every per-state Chern value integral
    within 1e-9 (on this 4 x 4 mesh the per-sample sums need not be 1: the plaquettes do not
    resolve every state's Berry curvature), every plaquette phase clear of +-pi."""
    from rlm_amd import HIP_DIAG_FP64
    print("[test_rlm_n3000_g4_fp64]", flush=True)
    e, c, m, _ = _run_n3000(HIP_DIAG_FP64)
    print(f"[test_rlm_n3000_g4_fp64] sums {np.rint(c).sum(axis=1)}, Chern dev {np.abs(c - np.rint(c)).max():.2e}, "
          f"min margin {m.min():.3e}", flush=True)
    assert np.abs(c - np.rint(c)).max() < 1e-9, np.abs(c - np.rint(c)).max()
    assert m.min() > 0.0
    assert np.all(np.diff(e, axis=1) >= 0)


def test_rlm_n3000_g4_mixed_vs_fp64():
    """Mixed plan at the n_phi = 3000 configuration against the FP64 plan: Chern integers
    identical, energies within 1e-10 relative. Plain FP32 vectors with FP64 Rayleigh
    quotients flip one state of sample 0 here (plaquette margins ~1e-3 rad) and leave
    3.5e-10 in the energies; the plan's FP64 eigenvector refinement (on for n_phi > 512)
    removes both."""
    from rlm_amd import HIP_DIAG_FP64, HIP_DIAG_MIXED
    print("[test_rlm_n3000_g4_mixed_vs_fp64]", flush=True)
    e64, c64, m64, _ = _run_n3000(HIP_DIAG_FP64)
    e32, c32, m32, res = _run_n3000(HIP_DIAG_MIXED)
    scale = np.abs(e64).max()
    nd = int((np.rint(c64) != np.rint(c32)).sum())
    print(f"[test_rlm_n3000_g4_mixed_vs_fp64] differing integers {nd}, energy rel diff "
          f"{np.abs(e32 - e64).max() / scale:.3e}, max residual {res:.3e} (|E| max {scale:.3e}), "
          f"mixed Chern dev {np.abs(c32 - np.rint(c32)).max():.2e}, min margin {m32.min():.3e}", flush=True)
    assert np.abs(c32 - np.rint(c32)).max() < 1e-9
    assert nd == 0
    assert np.abs(e32 - e64).max() <= 1e-10 * scale, np.abs(e32 - e64).max() / scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_split_levels_many_nodes(monkeypatch, precision):
    """n = 1030 with every merge level from halves of 32 rows up split across workgroups
    (RLM_DC_SPLIT_LEVEL=6: 17 nodes at the first split level, several rounds of nodes
    per workgroup before it stops) against LAPACK."""
    import torch
    from rlm_amd import heevd_batched, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print(f"[test_split_levels_many_nodes] {precision}", flush=True)
    monkeypatch.setenv("RLM_DC_SPLIT_LEVEL", "6")
    rng = np.random.default_rng(1030)
    H = _rand_herm(rng, 2, 1030)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    w, U = heevd_batched(torch.from_numpy(H).cuda(), precision=prec)
    (_check_mixed if precision == "mixed" else _check)(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_n3000_random_hermitian(precision):
    """One random 3000 x 3000 Hermitian matrix (the C5 size) against LAPACK."""
    import torch
    from rlm_amd import heevd_batched, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print(f"[test_n3000_random_hermitian] {precision}", flush=True)
    rng = np.random.default_rng(3000)
    H = _rand_herm(rng, 1, 3000)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    w, U = heevd_batched(torch.from_numpy(H).cuda(), precision=prec)
    (_check_mixed if precision == "mixed" else _check)(H, w.cpu().numpy(), U.cpu().numpy())


@pytest.mark.parametrize("n_phi,g,refine", [(600, 2, "0"), (256, 3, "1")])
def test_rlm_mixed_refinement_switch_vs_live_oracle(oracle, monkeypatch, n_phi, g, refine):
    """RLM_REFINE=0 (FP32 vectors with FP64 Rayleigh quotients at n_phi = 600) and
    RLM_REFINE=1 (the FP64 refinement at n_phi = 256, band-form reduction) against the
    CPU oracle: Chern integers identical, energies within 1e-10 relative."""
    from rlm_amd import RlmPlan, HIP_DIAG_MIXED
    print(f"[test_rlm_mixed_refinement_switch_vs_live_oracle] n={n_phi} g={g} RLM_REFINE={refine}", flush=True)
    monkeypatch.setenv("RLM_REFINE", refine)
    plan = RlmPlan(n_phi, 1.0, 16.0, g, 3, samples_per_batch=1, precision=HIP_DIAG_MIXED)
    assert plan.refine_info()[0] == (refine == "1")
    e, c, m = plan.run(0, 1)
    plan.close()
    r = oracle.rlm_sample(n_phi, 1.0, 16.0, g, 3, 0, threads=8)
    assert m.min() > 1e-3, m.min()
    _cmp_sample(e[0], c[0], r.energy, r.chern, (n_phi, g, "refine", refine))
