"""Pins the oracle's MT19937 stream (MKL VSL_BRNG_MT19937 restated) three ways:
the published mt19937ar known answer, CPython's independent `random` module, and
Intel MKL's own output (tests/golden/mkl_rng.npz, written by make_golden.py)."""
import json
import os
import random

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mkl_rng.npz")


def test_mt19937ar_known_answer(oracle):
    # mt19937ar.c (Matsumoto & Nishimura 2002) main(): init_by_array({0x123,0x234,0x345,0x456})
    # prints 1067595299 955945823 477289528 4107218783 4228976476 as its first outputs.
    import ctypes
    L = oracle.lib()
    L.orc_mt_state_size.restype = ctypes.c_size_t
    st = ctypes.create_string_buffer(L.orc_mt_state_size())
    key = (ctypes.c_uint32 * 4)(0x123, 0x234, 0x345, 0x456)
    L.orc_mt_init_by_array(st, key, 4)
    L.orc_mt_next.restype = ctypes.c_uint32
    got = [L.orc_mt_next(st) for _ in range(5)]
    assert got == [1067595299, 955945823, 477289528, 4107218783, 4228976476]


@pytest.mark.parametrize("seed", [0, 1, 3, 7, 11, 12345, 2**32 - 1])
def test_matches_cpython_random(oracle, seed):
    # CPython seeds its MT19937 with init_by_array over the 32-bit words of |seed|
    r = random.Random(seed)
    ref = np.array([r.getrandbits(32) for _ in range(2000)], dtype=np.uint32)
    assert np.array_equal(oracle.mt_words(seed, 0, 2000), ref)
    assert np.array_equal(oracle.mt_words(seed, 1234, 500), ref[1234:1734])


def test_rlm_disorder_matches_mkl(oracle):
    z = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(z["meta_json"]).decode())
    assert meta["rlm_cases"]
    for key, seed, n_imp, L, sample, n in meta["rlm_cases"]:
        sz = oracle.RlmSizes(0, L, L, n_imp, 0, 0)
        x, y, s = oracle.rlm_disorder(sz, seed, sample)
        assert np.array_equal(x[:n], z[key + "_x"]), key
        assert np.array_equal(y[:n], z[key + "_y"]), key
        bits = z[key + "_bits"]
        assert np.array_equal(s[:n], np.where(bits % 2 == 0, 1.0, -1.0)), key


def test_hlm_disorder_matches_mkl(oracle):
    z = np.load(GOLDEN, allow_pickle=False)
    for seed, ns, nsam in [(1, 576, 3), (5, 9, 4)]:
        assert np.array_equal(oracle.hlm_disorder(seed, ns, nsam), z[f"hlm_s{seed}_n{ns}"])


def test_uniform_mapping_edges(oracle):
    # u = a + (b-a) * w * 2^-32 on [0, L): w = 0 gives exactly 0, never L
    w = oracle.mt_words(1, 0, 10000).astype(np.float64)
    sz = oracle.RlmSizes(0, 4.0, 4.0, 10000, 0, 0)
    x, _, _ = oracle.rlm_disorder(sz, 1, 0)
    assert np.array_equal(x, 4.0 * (w * 2.0**-32))
    assert x.min() >= 0.0 and x.max() < 4.0


def test_fp32_build_disorder_matches_mkl(oracle):
    """The single-precision GPU builds' stream (dtype float: vsRngUniform, float L and
    impurity count) against MKL's own float output, including a non-power-of-two L
    (n_phi = 3000, 300): L * float(word) * 2^-32 in float arithmetic."""
    z = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(z["meta_json"]).decode())
    assert meta["rlm_f32_cases"]
    for key, seed, n_phi, n_imp, sample, n in meta["rlm_f32_cases"]:
        sz = oracle.rlm_sizes_f32(n_phi)
        assert sz.n_imp == n_imp
        x, y, _ = oracle.rlm_disorder_f32(sz, seed, sample)
        assert np.array_equal(x[:n].astype(np.float32), z[key + "_x"]), key
        assert np.array_equal(y[:n].astype(np.float32), z[key + "_y"]), key
