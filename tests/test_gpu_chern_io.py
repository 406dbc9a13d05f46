"""Chern drop-ins through the C ABI: hip_cal_chern (cal_Chern, diag_wrappers/chern.cpp:3-35)
on 2-slot rings of oracle eigenvectors (rlm dim_vec = n_phi, hlm dim_vec = n_sites, both
theta_1 parities), and the -DwfsIO path: rlm_plan_write_wfs / hip_write_wfs files
(diag_wrappers/wfs_file.cpp:2-23) read back by hip_cal_chern_wfs_io (chern.cpp:77-123),
checked against the oracle's chern_line on the same files and against the plan's own
per-state Chern numbers."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rlm_line(oracle, sz, V, cmt, cjm, t1, g):
    _, vecs = oracle._solve_line_rlm(sz, V, cmt, cjm, t1, g, None)
    return vecs                                   # [g+1, state, component]


@pytest.mark.parametrize("theta_1", [1, 2, 5, 6])
def test_cal_chern_rlm_vs_oracle(oracle, theta_1):
    from rlm_amd import cal_chern
    n, g = 16, 6
    print(f"[test_cal_chern_rlm_vs_oracle] theta_1={theta_1}", flush=True)
    sz = oracle.rlm_sizes(n)
    cmn, cmt, cjm = oracle.rlm_tables(sz, g)
    V = oracle.rlm_potential(sz, *oracle.rlm_disorder(sz, 1, 0), cmn)
    ring = np.zeros((2, g + 1, n, n), np.complex128)
    ring[(theta_1 - 1) % 2] = _rlm_line(oracle, sz, V, cmt, cjm, theta_1 - 1, g)
    ring[theta_1 % 2] = _rlm_line(oracle, sz, V, cmt, cjm, theta_1, g)
    ref = oracle.chern_line(ring, n, n, g, theta_1)
    got = cal_chern(ring, n, n, g, theta_1)
    assert np.abs(got - ref).max() < 1e-12, np.abs(got - ref).max()


@pytest.mark.parametrize("theta_1", [1, 4])
def test_cal_chern_hlm_vs_oracle(oracle, theta_1):
    """dim_vec = n_sites = 64 > n_phi = 16 (8x8 lattice, flux 1/4)."""
    from rlm_amd import cal_chern
    lx = ly = 8
    Q, band, g = 4, 0, 4
    n_sites, n = lx * ly, lx * ly // Q
    print(f"[test_cal_chern_hlm_vs_oracle] theta_1={theta_1}", flush=True)
    pot = oracle.hlm_disorder(3, n_sites, 1)[0]
    ring = np.zeros((2, g + 1, n, n_sites), np.complex128)
    for t1, slot in ((theta_1 - 1, (theta_1 - 1) % 2), (theta_1, theta_1 % 2)):
        for t2 in range(g):
            ring[slot, t2] = oracle.hlm_solve_twist(lx, ly, Q, band, g, pot, t1, t2)[1]
        ring[slot, g] = ring[slot, 0]             # hlm/main_hlm_cpu.cpp:117-118,149-150
    ref = oracle.chern_line(ring, n, n_sites, g, theta_1)
    got = cal_chern(ring, n, n_sites, g, theta_1)
    assert np.abs(got - ref).max() < 1e-12, np.abs(got - ref).max()


def _read_ring(directory, n, g, elem_bytes):
    dt = np.complex128 if elem_bytes == 16 else np.complex64
    ring = np.zeros((2, g + 1, n, n), np.complex128)
    for slot in range(2):
        for t2 in range(g + 1):
            v = np.fromfile(os.path.join(directory, f"wfs_{slot}_{t2}.bin"), dtype=dt)
            assert v.size == n * n
            ring[slot, t2] = v.reshape(n, n)
    return ring


@pytest.mark.parametrize("precision,elem_bytes", [("fp64", 16), ("mixed", 8), ("fp64", 8)])
def test_plan_wfs_dump_and_streamed_chern(tmp_path, oracle, precision, elem_bytes):
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED, cal_chern_wfs_io
    n, g, seed = 16, 6, 1
    print(f"[test_plan_wfs_dump_and_streamed_chern] {precision} elem={elem_bytes}", flush=True)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    plan = RlmPlan(n, 1.0, 16.0, g, seed, samples_per_batch=2, precision=prec)
    e, c, m = plan.run(0, 2)
    d = str(tmp_path)
    total = np.zeros(n)
    sample = 1
    plan.write_wfs(sample, 0, d, elem_bytes)
    for t1 in range(1, g + 1):
        plan.write_wfs(sample, t1, d, elem_bytes)
        ring = _read_ring(d, n, g, elem_bytes)
        ref = oracle.chern_line(ring, n, n, g, t1)
        got = cal_chern_wfs_io(n, n, g, t1, d, elem_bytes)
        assert np.abs(got - ref).max() < 1e-12, (t1, np.abs(got - ref).max())
        total += got.reshape(n, g).sum(axis=1)
    total += 1.0 / n
    # the dumped vectors are a gauge of the plan's own: the same per-state numbers
    # (complex64 files round the FP64 vectors to FP32: integers only)
    tol = 1e-9 if elem_bytes == 16 else 1e-3
    assert np.abs(total - c[sample]).max() < tol, np.abs(total - c[sample]).max()
    assert np.array_equal(np.rint(total), np.rint(c[sample]))
    plan.close()


def test_write_wfs_and_missing_file(tmp_path):
    import ctypes
    from rlm_amd import lib, cal_chern_wfs_io, HipDiagError
    rng = np.random.default_rng(5)
    a = rng.standard_normal(24) + 1j * rng.standard_normal(24)
    d = str(tmp_path)
    for eb, dt in ((16, np.complex128), (8, np.complex64)):
        rc = lib().hip_write_wfs(d.encode(), 1, 3, a.ctypes.data_as(ctypes.c_void_p), a.size, eb)
        assert rc == 0
        b = np.fromfile(os.path.join(d, "wfs_1_3.bin"), dtype=dt)
        assert np.array_equal(b, a.astype(dt))
    with pytest.raises(HipDiagError, match="status -2"):
        cal_chern_wfs_io(4, 4, 2, 1, os.path.join(d, "nowhere"))
