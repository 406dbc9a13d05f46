"""End-to-end invariants of the oracle's rlm/hlm drivers (rlm/main_rlm_cpu.cpp,
hlm/main_hlm_cpu.cpp): per-state Chern values are integers and sum to 1 per sample
(README.md:58,79), the rlm energies sum to the trace N*fl32(1/N)*sum(s_i)
(README.md:82-113 shows sums 0 and -20), and the committed fixtures reproduce."""
import numpy as np
import pytest


@pytest.mark.parametrize("n_phi,g,seed,sample", [(16, 8, 1, 0), (16, 10, 2, 3), (24, 6, 9, 1), (32, 4, 5, 0)])
def test_rlm_sample_invariants(oracle, n_phi, g, seed, sample):
    r = oracle.rlm_sample(n_phi, 1.0, 16.0, g, seed, sample)
    assert np.abs(r.chern - np.rint(r.chern)).max() < 1e-9
    assert np.rint(r.chern).sum() == 1
    sz = oracle.rlm_sizes(n_phi)
    _, _, s = oracle.rlm_disorder(sz, seed, sample)
    tr = n_phi * float(np.float32(1.0) / np.float32(n_phi)) * s.sum()
    assert abs(r.energy.sum() - tr) <= 1e-10 * max(1.0, abs(tr))
    assert np.all(np.diff(r.energy) >= 0)


def test_rlm_fixture_reproduces(oracle, fixtures):
    fx = fixtures["rlm_n16_g64_s1"]
    p = fx["params"]
    row = fx["samples"][1]
    r = oracle.rlm_sample(p["n_phi"], p["q"], p["imp"], p["n_mesh"], p["seed"], row["sample"], threads=4)
    assert np.abs(r.energy - np.array(row["energy"])).max() < 1e-12
    assert np.array_equal(np.rint(r.chern), np.rint(row["chern"]))


def test_fixture_sums(fixtures):
    for name, fx in fixtures.items():
        for row in fx["samples"]:
            c = np.array(row["chern"])
            assert np.abs(c - np.rint(c)).max() < 1e-8, name
            assert np.rint(c).sum() == 1, name


@pytest.mark.parametrize("lx,ly,Q,g,seed", [(6, 6, 3, 6, 1), (3, 3, 3, 8, 4)])
def test_hlm_sample_invariants(oracle, lx, ly, Q, g, seed):
    pots = oracle.hlm_disorder(seed, lx * ly, 2)
    for s in range(2):
        r = oracle.hlm_sample(lx, ly, Q, 0, g, pots[s])
        assert np.abs(r.chern - np.rint(r.chern)).max() < 1e-9
        assert np.rint(r.chern).sum() == 1
