"""GPU parity of the batched hlm pipeline (librlm_hip.so) against the CPU oracle
(hlm/main_hlm_cpu.cpp restated): Chern integers identical, sums 1, energies within
1e-10 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cmp(e, c, ref_e, ref_c, tag):
    ref_e = np.asarray(ref_e)
    scale = np.abs(ref_e).max()
    assert np.abs(e - ref_e).max() <= 1e-10 * scale, (tag, np.abs(e - ref_e).max() / scale)
    assert np.array_equal(np.rint(c), np.rint(ref_c)), (tag, np.rint(c), np.rint(ref_c))
    assert np.rint(c).sum() == 1, tag


@pytest.mark.parametrize("lx,ly,Q,g,seed,S", [(3, 3, 3, 8, 4, 2), (6, 6, 3, 6, 1, 2), (8, 4, 4, 5, 2, 1)])
def test_small_vs_live_oracle(oracle, lx, ly, Q, g, seed, S):
    from rlm_amd import HlmPlan
    print(f"[hlm small] {lx}x{ly} Q={Q} g={g} seed={seed}", flush=True)
    plan = HlmPlan(lx, ly, Q, 0, g, seed, 2.0, samples_per_batch=S)
    e, c, m = plan.run(0, S)
    pots = oracle.hlm_disorder(seed, lx * ly, S)
    for s in range(S):
        r = oracle.hlm_sample(lx, ly, Q, 0, g, pots[s])
        _cmp(e[s], c[s], r.energy, r.chern, (lx, ly, Q, s))
    plan.close()


def test_config_fixture(fixtures):
    """24x24, q=8, b=0, g=32, seed 1 (the batched hlm configuration at the reference's W=2)."""
    from rlm_amd import HlmPlan
    fx = fixtures["hlm_24x24_q8_b0_g32_s1"]
    p = fx["params"]
    print("[hlm fixture] 24x24 q=8 g=32", flush=True)
    plan = HlmPlan(p["lx"], p["ly"], p["Q"], p["band"], p["n_mesh"], p["seed"], 2.0, samples_per_batch=2)
    e, c, m = plan.run(0, 2)
    for k, row in enumerate(fx["samples"]):
        _cmp(e[k], c[k], row["energy"], row["chern"], ("hlm fixture", k))
    plan.close()
