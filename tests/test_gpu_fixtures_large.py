"""GPU plans against CPU-oracle fixtures at the full configuration shapes
(tests/golden/oracle_fixtures_large.json, written by tests/golden/make_golden.py large):

* rlm n_phi = 3000, g = 4, seed 3, samples 0-1 (-q 1 -i 16: 3000 x 3000 matrices), FP64 and
  mixed precision with FP64 refinement;
* the uneven-batch sizes n_phi = 192, 256, 384, 512 at g = 16, seed 11, sample 0, FP64 and mixed;
* hlm 24x24, q = 8, b = 0, g = 32, seed 1 at disorder width W = 1, samples 0-3.

Bar: every per-state Chern number rounded to an integer identical to the oracle's, the
per-sample sums equal, mesh-averaged energies within 1e-10 relative; each state's
deviation from the nearest integer (plan and oracle) and from the oracle's value is
printed and bounded."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E_RTOL = 1e-10


@pytest.fixture(scope="module")
def large():
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_fixtures_large.json")) as f:
        return json.load(f)


def _cmp(e, c, row, tag, dev_tol):
    ref_e = np.asarray(row["energy"])
    ref_c = np.asarray(row["chern"])
    scale = np.abs(ref_e).max()
    e_rel = np.abs(e - ref_e).max() / scale
    dev = np.abs(c - np.rint(c)).max()
    vs = np.abs(c - ref_c).max()
    print(f"[fixture] {tag}: energy rel diff {e_rel:.2e}, plan max |C - round C| {dev:.2e}, oracle "
          f"{row['max_int_dev']:.2e}, max |C_plan - C_oracle| {vs:.2e}, sum {np.rint(c).sum():.0f}", flush=True)
    assert e_rel <= E_RTOL, (tag, e_rel)
    assert np.array_equal(np.rint(c), np.rint(ref_c)), (tag, np.nonzero(np.rint(c) != np.rint(ref_c)))
    assert np.rint(c).sum() == np.rint(ref_c).sum(), tag
    assert dev < dev_tol, (tag, dev)
    assert vs < dev_tol + row["max_int_dev"], (tag, vs)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["rlm_n192_g16_s11", "rlm_n256_g16_s11", "rlm_n384_g16_s11", "rlm_n512_g16_s11"])
def test_uneven_sizes_g16(large, name, precision):
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    fx = large[name]
    p = fx["params"]
    print(f"[test_uneven_sizes_g16] {name} {precision}", flush=True)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    plan = RlmPlan(p["n_phi"], p["q"], p["imp"], p["n_mesh"], p["seed"], samples_per_batch=1, precision=prec)
    e, c, m = plan.run(0, 1)
    plan.close()
    _cmp(e[0], c[0], fx["samples"][0], (name, precision), 1e-9)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rlm_n3000_g4_vs_oracle(large, precision):
    """The n_phi = 3000 configuration, samples 0 and 1 (25 solves of 3000 x 3000 each in
    the oracle). Mixed: FP32 reduction and back-transform, FP64 refinement (on for
    n_phi > 512)."""
    from rlm_amd import RlmPlan, HIP_DIAG_FP64, HIP_DIAG_MIXED
    fx = large["rlm_n3000_g4_s3"]
    p = fx["params"]
    print(f"[test_rlm_n3000_g4_vs_oracle] {precision}", flush=True)
    prec = HIP_DIAG_MIXED if precision == "mixed" else HIP_DIAG_FP64
    samples = [r["sample"] for r in fx["samples"]]
    plan = RlmPlan(p["n_phi"], p["q"], p["imp"], p["n_mesh"], p["seed"], samples_per_batch=len(samples),
                   precision=prec)
    e, c, m = plan.run(samples[0], len(samples))
    plan.close()
    print(f"[test_rlm_n3000_g4_vs_oracle] min plaquette margin {m.min():.3e} rad", flush=True)
    for k, row in enumerate(fx["samples"]):
        _cmp(e[k], c[k], row, ("n3000", precision, row["sample"]), 1e-9)


def test_hlm_c2_w1_vs_oracle(large):
    from rlm_amd import HlmPlan
    fx = large["hlm_24x24_q8_b0_g32_s1_w1"]
    p = fx["params"]
    samples = [r["sample"] for r in fx["samples"]]
    print(f"[test_hlm_c2_w1_vs_oracle] samples {samples}", flush=True)
    plan = HlmPlan(p["lx"], p["ly"], p["Q"], p["band"], p["n_mesh"], p["seed"], p["width"],
                   samples_per_batch=len(samples))
    e, c, m = plan.run(samples[0], len(samples))
    plan.close()
    for k, row in enumerate(fx["samples"]):
        # the hlm plans solve every point of the twist mesh (the reference's theta_2 = g copy
        # included), so each value carries the oracle's own ~1e-13 deviation from an integer
        _cmp(e[k], c[k], row, ("hlm W=1", row["sample"]), 1e-9)
