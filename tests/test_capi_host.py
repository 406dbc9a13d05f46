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
