"""The single-precision GPU builds' conventions (HIP_DIAG_CONV_GPU_FP32, bin/rlm -G):
float sizes, tables and impurity positions, V_mn with e^{+i phase}
(rlm/potential_coeff.cu:14-15), h_jk with conj(c1) and c2 (rlm/hamiltonian_rlm_gpu.cu:23,36),
every twist point solved, energy_Chern_No.dat rows (rlm/main_rlm_gpu.cu:296-297). Against
the oracle's FP64 evaluation of the same formulas on the same float inputs: energies within
1e-10 relative, integers identical, Chern values within 1e-9 (FP64 plans) or 1e-4 (mixed
plans: with every twist point solved the plaquette sums do not telescope, so the FP32
eigenvectors' ~1e-7 errors stay in the values, as in the FP32 builds)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "random-landau-model_amd", "bin")


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n_phi,g,seed,S", [(16, 8, 1, 2), (64, 6, 11, 1), (300, 3, 11, 1)])
def test_fp32_build_conventions_vs_oracle(oracle, n_phi, g, seed, S, precision):
    from rlm_amd import RlmPlan, HIP_DIAG_CONV_GPU_FP32
    print(f"[test_fp32_build_conventions_vs_oracle] n={n_phi} g={g} seed={seed} precision {precision}", flush=True)
    plan = RlmPlan(n_phi, 1.0, 16.0, g, seed, samples_per_batch=S, precision=precision, flags=HIP_DIAG_CONV_GPU_FP32)
    assert plan.n_imp == oracle.rlm_sizes_f32(n_phi).n_imp
    e, c, m = plan.run(0, S)
    plan.close()
    for s in range(S):
        r = oracle.rlm_sample_gpu_conv(n_phi, 1.0, 16.0, g, seed, s, threads=8)
        scale = np.abs(r.energy).max()
        assert np.abs(e[s] - r.energy).max() <= 1e-10 * scale, np.abs(e[s] - r.energy).max() / scale
        assert np.abs(c[s] - r.chern).max() < (1e-9 if precision == 0 else 1e-4), np.abs(c[s] - r.chern).max()
        assert np.array_equal(np.rint(c[s]), np.rint(r.chern))


def test_rlm_cli_gpu_build_rows(tmp_path, oracle):
    print("[test_rlm_cli_gpu_build_rows]", flush=True)
    r = subprocess.run([os.path.join(BIN, "rlm"), "-n", "16", "-g", "8", "-c", "2", "-s", "1", "-G"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "energy_Chern_No.dat").read_text().splitlines()
    assert len(lines) == 32 and not (tmp_path / "energy_Chern_No.txt").exists()
    for s in range(2):
        ref = oracle.rlm_sample_gpu_conv(16, 1.0, 16.0, 8, 1, s)
        got = [l for l in lines if int(l.split()[0]) == s]
        for l, ev in zip(got, ref.energy):
            assert l[:9] == f"{s:>8} " and l[24] == " ", l          # setw(8), setw(15), right-aligned
            assert abs(float(l[9:24]) - ev) <= 1e-6 * np.abs(ref.energy).max()
        c = np.array([float(l.split()[2]) for l in got])
        assert np.array_equal(np.rint(c), np.rint(ref.chern))
