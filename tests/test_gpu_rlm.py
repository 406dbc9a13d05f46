"""GPU parity of the batched rlm pipeline (librlm_hip.so) against the CPU oracle.

Bar: every per-state Chern number rounded to an integer
identical to the oracle's, every per-sample sum = 1, mesh-averaged energies within
1e-10 relative. Intermediate stages (V_mn, assembled H) are compared elementwise
with rtol 1e-12 (FP64 sums in a different order).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E_RTOL = 1e-10


def _cmp_sample(e, c, ref_e, ref_c, tag, int_tol=1e-9):
    """int_tol bounds |C - round(C)| per state. The plans derive the theta = 2 pi edges of
    the twist mesh from the theta = 0 vectors by the exact boundary relation, so the
    plaquette sum telescopes to an integer up to the rounding of the atan2 sums (~1e-15)
    in both precisions; the reference solves the edges separately and leaves ~1e-13
    (FP64) or ~1e-5 (FP32 vectors)."""
    ref_e = np.asarray(ref_e)
    ref_c = np.asarray(ref_c)
    scale = np.abs(ref_e).max()
    assert np.abs(e - ref_e).max() <= E_RTOL * scale, (tag, np.abs(e - ref_e).max() / scale)
    assert np.array_equal(np.rint(c), np.rint(ref_c)), (tag, np.rint(c), np.rint(ref_c))
    # the per-sample sum is 1 once the mesh resolves every plaquette phase; on coarse
    # meshes (n_phi=8, g=6) the reference's own sum can differ, and must be matched
    assert np.rint(c).sum() == np.rint(ref_c).sum(), tag
    assert np.abs(c - np.rint(c)).max() < int_tol, (tag, np.abs(c - np.rint(c)).max())


@pytest.mark.parametrize("n_phi,g,seed,samples", [(16, 8, 1, 2), (16, 5, 3, 1), (8, 6, 2, 3), (24, 4, 5, 1)])
def test_small_vs_live_oracle(oracle, n_phi, g, seed, samples):
    from rlm_amd import RlmPlan
    print(f"[test_small_vs_live_oracle] n={n_phi} g={g} seed={seed} samples={samples}", flush=True)
    plan = RlmPlan(n_phi, 1.0, 16.0, g, seed, samples_per_batch=samples)
    e, c, m = plan.run(0, samples)
    for s in range(samples):
        r = oracle.rlm_sample(n_phi, 1.0, 16.0, g, seed, s)
        _cmp_sample(e[s], c[s], r.energy, r.chern, (n_phi, g, seed, s))
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n_phi,g,seed", [(192, 4, 11), (384, 3, 11), (512, 2, 11)])
def test_uneven_sizes_vs_live_oracle(oracle, n_phi, g, seed, precision):
    """The uneven-batch sizes beyond the register-resident kernels' range (n_phi = 192, 384,
    512; -q 1 -i 16, seed 11, sample 0) on a coarse mesh, FP64 and mixed, against the oracle:
    dense-matrix fallbacks of the band-form plans, the lower-storage reductions and the VALU
    back-transform (n > 256)."""
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print(f"[test_uneven_sizes_vs_live_oracle] n={n_phi} g={g} {precision}", flush=True)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    plan = RlmPlan(n_phi, 1.0, 16.0, g, seed, samples_per_batch=1, precision=prec)
    e, c, m = plan.run(0, 1)
    r = oracle.rlm_sample(n_phi, 1.0, 16.0, g, seed, 0, threads=8)
    # the integer comparison is meaningful while every plaquette phase stays clear of +-pi
    assert m.min() > 1e-3, m.min()
    _cmp_sample(e[0], c[0], r.energy, r.chern, (n_phi, g, precision))
    plan.close()


@pytest.mark.parametrize("name", ["rlm_n16_g64_s1", "rlm_n64_g16_s11", "rlm_n128_g16_s11", "rlm_n256_g32_s7"])
def test_config_fixtures(fixtures, name):
    from rlm_amd import RlmPlan
    fx = fixtures[name]
    p = fx["params"]
    samples = [r["sample"] for r in fx["samples"]]
    print(f"[test_config_fixtures] {name} samples={samples}", flush=True)
    plan = RlmPlan(p["n_phi"], p["q"], p["imp"], p["n_mesh"], p["seed"], samples_per_batch=len(samples))
    e, c, m = plan.run(samples[0], len(samples))
    for k, row in enumerate(fx["samples"]):
        _cmp_sample(e[k], c[k], row["energy"], row["chern"], (name, row["sample"]))
        assert np.rint(c[k]).sum() == 1
    plan.close()


def test_potential_and_assembly_vs_oracle(oracle):
    """hip_prepare_potential_coeff and hip_set_hamil_matrix against the oracle's V_mn and H."""
    import torch
    from rlm_amd import lib, prepare_coeff, generate_disorder
    print("[test_potential_and_assembly_vs_oracle]", flush=True)
    for n_phi, g, seed, sample, t1, t2 in [(16, 8, 1, 1, 3, 5), (64, 16, 11, 0, 0, 16), (256, 32, 7, 1, 17, 9)]:
        sz = oracle.rlm_sizes(n_phi)
        cmn, cmt, cjm = oracle.rlm_tables(sz, g)
        x, y, s = oracle.rlm_disorder(sz, seed, sample)
        V = oracle.rlm_potential(sz, x, y, s, cmn)
        H = oracle.rlm_hamiltonian(sz, V, cmt, cjm, t1, t2)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        dx, dy, ds, dcmn = dev(x), dev(y), dev(s), dev(cmn)
        dV = torch.zeros((sz.d, sz.d), dtype=torch.complex128, device="cuda")
        rc = lib().hip_prepare_potential_coeff(dx.data_ptr(), dy.data_ptr(), ds.data_ptr(), dV.data_ptr(),
                                               dcmn.data_ptr(), sz.d, sz.d, sz.off, sz.n_imp)
        assert rc == 0
        Vg = dV.cpu().numpy()
        assert np.abs(Vg - V).max() <= 1e-12 * np.abs(V).max(), np.abs(Vg - V).max() / np.abs(V).max()
        dc1, dc2, dcjm = dev(cmt[t1]), dev(cmt[t2]), dev(cjm)
        dVo = dev(V)
        dA = torch.zeros((n_phi, n_phi), dtype=torch.complex128, device="cuda")
        rc = lib().hip_set_hamil_matrix(dc1.data_ptr(), dc2.data_ptr(), dcjm.data_ptr(), dVo.data_ptr(),
                                        dA.data_ptr(), n_phi, sz.d, sz.off)
        assert rc == 0
        Hg = dA.cpu().numpy().T     # column-major buffer -> matrix
        err = np.abs(Hg - H).max() / np.abs(H).max()
        assert err <= 1e-12, err


def test_full_size_properties():
    """n_phi=256, g=32 (1089 twist points), 4 samples of seed 7: integral Chern numbers,
    per-sample sum 1, energies sorted per sample mean, and sum of energies = trace invariant."""
    from rlm_amd import RlmPlan
    print("[test_full_size_properties]", flush=True)
    plan = RlmPlan(256, 1.0, 16.0, 32, 7, samples_per_batch=4)
    e, c, m = plan.run(4, 4)
    st = plan.stage_times()
    assert list(st) == list(RlmPlan.STAGES) and all(v >= 0 for v in st.values()), st
    assert np.all(np.abs(c - np.rint(c)) < 1e-9)
    assert np.all(np.rint(c).sum(axis=1) == 1)
    assert np.all(m > 0)
    plan.close()


@pytest.mark.parametrize("name", ["rlm_n16_g64_s1", "rlm_n64_g16_s11", "rlm_n128_g16_s11", "rlm_n256_g32_s7"])
def test_mixed_precision_fixtures(fixtures, name):
    """Mixed path (FP32 tridiagonalisation/back-transform, FP64 D&C, FP64 Rayleigh
    quotients, FP64-accumulated links): same bar as FP64 -- Chern integers identical,
    energies within 1e-10 relative of the FP64 oracle."""
    from rlm_amd import RlmPlan, HIP_DIAG_MIXED
    fx = fixtures[name]
    p = fx["params"]
    samples = [r["sample"] for r in fx["samples"]]
    print(f"[test_mixed_precision_fixtures] {name}", flush=True)
    plan = RlmPlan(p["n_phi"], p["q"], p["imp"], p["n_mesh"], p["seed"], samples_per_batch=len(samples),
                   precision=HIP_DIAG_MIXED)
    e, c, m = plan.run(samples[0], len(samples))
    res = plan.max_residual()
    print(f"  max FP64 residual of FP32 eigenvectors: {res:.3e}; min plaquette margin {m.min():.3e}", flush=True)
    for k, row in enumerate(fx["samples"]):
        _cmp_sample(e[k], c[k], row["energy"], row["chern"], (name, "mixed", row["sample"]))
    assert res < 1e-3
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n_phi,g,seed", [(16, 1, 1), (64, 8, 11), (256, 4, 7)])
def test_derived_edges_match_full_mesh(monkeypatch, n_phi, g, seed, precision):
    """Plans diagonalise the g^2 points theta_1, theta_2 < g and derive the theta = 2 pi
    edges (u(t1, g)[i] = u(t1, 0)[(i+1) % N], u(g, t2)[i] = e^{i 2 pi i/N} u(0, t2)[i]);
    RLM_MESH=full solves all (g+1)^2 points as the reference does. Same inputs, both
    ways: energies within 1e-12 (FP64) / 1e-10 (mixed) relative, identical Chern integers, Chern values within
    1e-9 (FP64) / 1e-3 (mixed: the separately solved FP32 edges leave ~1e-4 at n = 256)
    of each other, and the derived-edge values integral to 1e-9 (the plaquette sum
    telescopes). g = 1 derives three of the four points of each sample."""
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print(f"[test_derived_edges_match_full_mesh] n={n_phi} g={g} seed={seed} {precision}", flush=True)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    plan = RlmPlan(n_phi, 1.0, 16.0, g, seed, samples_per_batch=2, precision=prec)
    e, c, m = plan.run(0, 2)
    plan.close()
    monkeypatch.setenv("RLM_MESH", "full")
    plan = RlmPlan(n_phi, 1.0, 16.0, g, seed, samples_per_batch=2, precision=prec)
    ef, cf, mf = plan.run(0, 2)
    plan.close()
    # mixed: Rayleigh quotients of two different FP32 vector sets (error ~ residual^2)
    e_tol = 1e-12 if prec == HIP_DIAG_FP64 else 1e-10
    assert np.abs(e - ef).max() <= e_tol * np.abs(ef).max(), np.abs(e - ef).max() / np.abs(ef).max()
    assert np.array_equal(np.rint(c), np.rint(cf))
    tol = 1e-9 if prec == HIP_DIAG_FP64 else 1e-3
    assert np.abs(c - cf).max() < tol, np.abs(c - cf).max()
    assert np.abs(c - np.rint(c)).max() < 1e-9, np.abs(c - np.rint(c)).max()
    print(f"  max |C - C_full| {np.abs(c - cf).max():.3e}; max |C - rint C| derived "
          f"{np.abs(c - np.rint(c)).max():.3e}, full {np.abs(cf - np.rint(cf)).max():.3e}", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n_phi,g,seed,S", [(64, 5, 11, 2), (300, 7, 3, 1), (64, 30, 11, 16)])
def test_tiled_links_identical_to_per_point(monkeypatch, n_phi, g, seed, S, precision):
    """links_tile_kernel reads the vector sets of a tile of B x B twist points (3 x 3; 4 x 4
    for FP64 launches of at least four workgroups per CU, which n_phi = 64, g = 30, 16
    samples reaches) and of its neighbours once per component row. Each link sums the
    components in the same order as the one-point-per-workgroup kernel (RLM_LINKS_TILE=0),
    so energies, Chern values and margins are bit-identical. The meshes of g + 1 = 6, 8 and
    31 points leave partial tiles at the theta = 2 pi edges; n_phi = 300 takes two chunks
    of 256 states."""
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print(f"[test_tiled_links_identical_to_per_point] n={n_phi} g={g} seed={seed} S={S} {precision}", flush=True)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    out = []
    for tile in ("", "0"):
        if tile:
            monkeypatch.setenv("RLM_LINKS_TILE", tile)
        plan = RlmPlan(n_phi, 1.0, 16.0, g, seed, samples_per_batch=S, precision=prec)
        out.append(plan.run(0, S))
        plan.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b), np.abs(a - b).max()


def test_mixed_pair_info_reports_runs():
    """rlm_plan_pair_info after a mixed execute at n_phi = 256 (seed 11 holds a
    near-degenerate pair at every twist point): runs found and solved, none dropped or
    left unrotated; FP64 plans report zeros."""
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print("[test_mixed_pair_info_reports_runs]", flush=True)
    plan = RlmPlan(256, 1.0, 16.0, 4, 11, samples_per_batch=1, precision=HIP_DIAG_MIXED)
    plan.run(0, 1)
    info = plan.pair_info()
    plan.close()
    assert info["solved"] >= 16 and info["dropped"] == 0 and info["unrotated"] == 0, info
    plan = RlmPlan(64, 1.0, 16.0, 4, 11, samples_per_batch=1, precision=HIP_DIAG_FP64)
    plan.run(0, 1)
    assert plan.pair_info() == {"solved": 0, "split": 0, "dropped": 0, "unrotated": 0}
    plan.close()


def test_mixed_near_degenerate_pair_vs_fp64():
    """The uneven-batch size n_phi = 256 at g = 16, 16 samples of seed 11: sample 0 holds a
    pair of states 2e-5 apart (states 15 and 16 at every twist point) whose FP32 vectors mix;
    plain Rayleigh quotients leave 1.3e-10 relative in the mesh-averaged energies there. The
    mixed plan's exact Rayleigh-Ritz of near-degenerate runs must bring the energies within
    1e-10 of the FP64 plan, with identical Chern integers."""
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    print("[test_mixed_near_degenerate_pair_vs_fp64] n=256 g=16 seed 11 x 16", flush=True)
    out = {}
    for prec in (HIP_DIAG_FP64, HIP_DIAG_MIXED):
        plan = RlmPlan(256, 1.0, 16.0, 16, 11, samples_per_batch=16, precision=prec)
        out[prec] = plan.run(0, 16)
        plan.close()
    e64, c64, _ = out[HIP_DIAG_FP64]
    e32, c32, _ = out[HIP_DIAG_MIXED]
    scale = np.abs(e64).max()
    assert np.abs(e32 - e64).max() <= E_RTOL * scale, np.abs(e32 - e64).max() / scale
    assert np.array_equal(np.rint(c32), np.rint(c64))
    assert np.abs(c32 - np.rint(c32)).max() < 1e-9
