"""Device MT19937 streams (hip_generate_disorder_potential, hip_hlm_generate_disorder)
against Intel MKL's vectors (tests/golden/mkl_rng.npz) and the library's host stream,
bit for bit. The plans generate their disorder with the same kernel at every execute,
so every plan test against the oracle also covers it."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mkl_rng.npz")


def test_device_rlm_streams_match_mkl():
    from rlm_amd import device_disorder
    z = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(z["meta_json"]).decode())
    for key, seed, n_imp, L, sample, n in meta["rlm_cases"]:
        print(f"[test_device_rlm_streams_match_mkl] {key}", flush=True)
        x, y, s = device_disorder(n_imp, L, L, seed, sample, 1)
        assert np.array_equal(x[0, :n], z[key + "_x"]), key
        assert np.array_equal(y[0, :n], z[key + "_y"]), key
        assert np.array_equal(s[0, :n], np.where(z[key + "_bits"] % 2 == 0, 1.0, -1.0)), key


@pytest.mark.parametrize("n_imp,seed,first,count", [(256, 1, 0, 3), (4096, 7, 0, 32), (4096, 7, 29, 5),
                                                    (48000, 3, 0, 4), (16, 5, 1000, 7), (1024, 11, 1 << 20, 2)])
def test_device_rlm_streams_match_host(n_imp, seed, first, count):
    """Every jump kind: none (sample 0), stepped (skip < 19937 words) and polynomial,
    as the batch's first-sample jump and as the per-sample jump."""
    from rlm_amd import device_disorder, generate_disorder
    print(f"[test_device_rlm_streams_match_host] n_imp={n_imp} seed={seed} first={first} count={count}", flush=True)
    L = float(np.sqrt(n_imp / 16.0))
    x, y, s = device_disorder(n_imp, L, L, seed, first, count)
    for k in range(count):
        hx, hy, hs = generate_disorder(n_imp, L, L, seed, first + k)
        assert np.array_equal(x[k], hx), k
        assert np.array_equal(y[k], hy), k
        assert np.array_equal(s[k], hs), k


def test_device_hlm_streams():
    from rlm_amd import device_hlm_disorder, hlm_disorder
    z = np.load(GOLDEN, allow_pickle=False)
    for seed, ns, nsam in [(1, 576, 3), (5, 9, 4)]:
        print(f"[test_device_hlm_streams] seed={seed} n_sites={ns}", flush=True)
        assert np.array_equal(device_hlm_disorder(ns, nsam, seed), z[f"hlm_s{seed}_n{ns}"])
    # samples 40..63 of the C2 stream (jumps of 23040 words and more), W = 1
    assert np.array_equal(device_hlm_disorder(576, 24, 1, 40, 1.0), hlm_disorder(576, 24, 1, 40, 1.0))
