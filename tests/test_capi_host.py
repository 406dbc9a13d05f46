"""The C-ABI library loads, exports every entry point include/hip_diag.h declares, and
its host-only entry points (tables, disorder stream) match the oracle bit for bit.
No GPU compute is called here."""
import numpy as np
import pytest


def test_library_exports_header_symbols():
    import ctypes
    from rlm_amd import LIB_PATH, header_symbols
    names = header_symbols()
    assert "hip_diag_hamil" in names and "rlm_plan_execute" in names
    so = ctypes.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


@pytest.mark.parametrize("n_phi,g", [(16, 64), (256, 32), (192, 16), (3000, 4)])
def test_prepare_coeff_matches_oracle(oracle, n_phi, g):
    from rlm_amd import prepare_coeff
    sz = oracle.rlm_sizes(n_phi)
    a = oracle.rlm_tables(sz, g)
    b = prepare_coeff(n_phi, sz.off, sz.d, g, sz.L1, sz.L2)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("n_phi,seed,sample", [(16, 1, 0), (16, 1, 1), (256, 7, 31), (3000, 3, 3)])
def test_disorder_matches_oracle(oracle, n_phi, seed, sample):
    from rlm_amd import generate_disorder
    sz = oracle.rlm_sizes(n_phi)
    a = oracle.rlm_disorder(sz, seed, sample)
    b = generate_disorder(sz.n_imp, sz.L1, sz.L2, seed, sample)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_hlm_disorder_matches_oracle(oracle):
    """hlm/main_hlm_cpu.cpp:66-67,79 single stream; W=1 is the W=2 stream halved exactly."""
    from rlm_amd import hlm_disorder
    ref = oracle.hlm_disorder(1, 576, 5)
    assert np.array_equal(hlm_disorder(576, 5, 1, 0, 2.0), ref)
    assert np.array_equal(hlm_disorder(576, 2, 1, 3, 2.0), ref[3:5])
    assert np.array_equal(hlm_disorder(576, 5, 1, 0, 1.0), 0.5 * ref)


def test_disorder_matches_mkl_vectors():
    """The library's host stream (jump-ahead by t^J mod phi) against Intel MKL's own
    vslSkipAheadStream output for every committed case (seed 3, n_imp 48000, sample 1
    skips 144000 words)."""
    import json
    import os
    from rlm_amd import generate_disorder
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "mkl_rng.npz"), allow_pickle=False)
    meta = json.loads(bytes(z["meta_json"]).decode())
    for key, seed, n_imp, L, sample, n in meta["rlm_cases"]:
        x, y, s = generate_disorder(n_imp, L, L, seed, sample)
        assert np.array_equal(x[:n], z[key + "_x"]), key
        assert np.array_equal(y[:n], z[key + "_y"]), key
        assert np.array_equal(s[:n], np.where(z[key + "_bits"] % 2 == 0, 1.0, -1.0)), key
    from rlm_amd import hlm_disorder
    for seed, ns, nsam in [(1, 576, 3), (5, 9, 4)]:
        assert np.array_equal(hlm_disorder(ns, nsam, seed), z[f"hlm_s{seed}_n{ns}"])


@pytest.mark.parametrize("n_imp,seed,sample", [(16, 1, 1000003), (48000, 3, 97)])
def test_long_jumps_match_stepping(oracle, n_imp, seed, sample):
    """Jumps of 4.8e7 and 1.4e7 words (polynomial path) against the oracle, which steps
    the generator word by word."""
    from rlm_amd import generate_disorder
    sz = oracle.RlmSizes(0, 1.0, 1.0, n_imp, 0, 0)
    ref = oracle.rlm_disorder(sz, seed, sample)
    got = generate_disorder(n_imp, 1.0, 1.0, seed, sample)
    for u, v in zip(ref, got):
        assert np.array_equal(u, v)


def _raw_words(n_imp, seed, sample):
    """Words 3 n_imp sample .. + 2 n_imp of the stream: x and y with L = 2^32 are the
    raw words themselves (w * 2^-32 * 2^32 is exact)."""
    from rlm_amd import generate_disorder
    x, y, _ = generate_disorder(n_imp, 2.0**32, 2.0**32, seed, sample)
    return np.concatenate([x, y])


@pytest.mark.parametrize("a,seed,sample", [(4096, 7, 1 << 20), (48000, 3, 123457)])
def test_jump_polynomials_agree(a, seed, sample):
    """Jumps far beyond what stepping can check (1.3e10 and 1.8e10 words): the stream
    reached by t^J1 mod phi must continue the one reached by a different polynomial
    t^J2 mod phi, J2 = J1 - d, after d stepped words."""
    J1 = 3 * a * sample
    w1 = _raw_words(a, seed, sample)
    for a2 in range(a + 1, a + 4000):
        r = J1 // (3 * a2)
        d = J1 - 3 * a2 * r
        if d + 64 <= a2:
            break
    w2 = _raw_words(a2, seed, r)
    n = a2 - d
    assert np.array_equal(w2[d:a2], w1[:n])
