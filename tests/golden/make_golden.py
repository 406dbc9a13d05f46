"""Writes the committed golden fixtures under tests/golden/.

1. ``mkl_rng.npz`` -- known-answer vectors of Intel MKL's VSL MT19937 stream,
   produced by calling MKL's own exported functions (vslNewStream,
   vslSkipAheadStream, vdRngUniform, vsRngUniform, viRngUniformBits) through
   ctypes, in the exact call sequences of rlm/disorder_potential.cpp:5-16 (double
   builds, and float builds for the single-precision GPU executables) and
   hlm/main_hlm_cpu.cpp:67,79. MKL is the third-party library the reference
   links (MKLROOT, Makefile:13,21); the runtime found here is
   /opt/conda/lib/libmkl_rt.so.1 (version printed into the file).
2. ``oracle_*.json`` -- end-to-end outputs (mesh-averaged energies, per-state
   Chern values) of the CPU oracle at the configuration shapes, so GPU parity
   tests need not rerun the slow CPU path.

Run from the repo root:  OPENBLAS_NUM_THREADS=1 python tests/golden/make_golden.py [mkl] [oracle]
"""
from __future__ import annotations

import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

MKL_PATH = "/opt/conda/lib/libmkl_rt.so.1"
VSL_BRNG_MT19937 = 8 << 20          # VSL_BRNG_SHIFT=20, MT19937 is generator #8
METHOD_STD = 0                      # VSL_RNG_METHOD_UNIFORM_STD / UNIFORMBITS_STD


def _mkl():
    mkl = ctypes.CDLL(MKL_PATH)
    vp = ctypes.c_void_p
    mkl.vslNewStream.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_uint]
    mkl.vslSkipAheadStream.argtypes = [vp, ctypes.c_longlong]
    mkl.vdRngUniform.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                 ctypes.c_double, ctypes.c_double]
    mkl.vsRngUniform.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                 ctypes.c_float, ctypes.c_float]
    mkl.viRngUniformBits.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint)]
    mkl.vslDeleteStream.argtypes = [ctypes.POINTER(vp)]
    return mkl


def _mkl_version(mkl) -> str:
    try:
        buf = ctypes.create_string_buffer(256)
        mkl.mkl_get_version_string.argtypes = [ctypes.c_char_p, ctypes.c_int]
        mkl.mkl_get_version_string(buf, 256)
        return buf.value.decode(errors="replace").strip()
    except Exception:   # pragma: no cover
        return "unknown"


def rlm_stream(mkl, seed, skip, n, L):
    """rlm/disorder_potential.cpp:5-16: x, y ~ U[0,L), raw bits."""
    st = ctypes.c_void_p()
    assert mkl.vslNewStream(ctypes.byref(st), VSL_BRNG_MT19937, seed) == 0
    assert mkl.vslSkipAheadStream(st, skip) == 0
    x = np.empty(n)
    y = np.empty(n)
    b = np.empty(n, dtype=np.uint32)
    dp = ctypes.POINTER(ctypes.c_double)
    assert mkl.vdRngUniform(METHOD_STD, st, n, x.ctypes.data_as(dp), 0.0, L) == 0
    assert mkl.vdRngUniform(METHOD_STD, st, n, y.ctypes.data_as(dp), 0.0, L) == 0
    assert mkl.viRngUniformBits(METHOD_STD, st, n, b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))) == 0
    mkl.vslDeleteStream(ctypes.byref(st))
    return x, y, b


def rlm_stream_f32(mkl, seed, skip, n, L):
    """rlm/disorder_potential.cpp:5-16 in the FP32 builds (dtype float: vsRngUniform)."""
    st = ctypes.c_void_p()
    assert mkl.vslNewStream(ctypes.byref(st), VSL_BRNG_MT19937, seed) == 0
    assert mkl.vslSkipAheadStream(st, skip) == 0
    x = np.empty(n, dtype=np.float32)
    y = np.empty(n, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    assert mkl.vsRngUniform(METHOD_STD, st, n, x.ctypes.data_as(fp), 0.0, float(np.float32(L))) == 0
    assert mkl.vsRngUniform(METHOD_STD, st, n, y.ctypes.data_as(fp), 0.0, float(np.float32(L))) == 0
    mkl.vslDeleteStream(ctypes.byref(st))
    return x, y


def hlm_stream(mkl, seed, n_sites, n_samples):
    """hlm/main_hlm_cpu.cpp:67,79: one stream, n_sites draws of U[-1,1) per sample."""
    st = ctypes.c_void_p()
    assert mkl.vslNewStream(ctypes.byref(st), VSL_BRNG_MT19937, seed) == 0
    out = np.empty((n_samples, n_sites))
    dp = ctypes.POINTER(ctypes.c_double)
    for s in range(n_samples):
        row = np.empty(n_sites)
        assert mkl.vdRngUniform(METHOD_STD, st, n_sites, row.ctypes.data_as(dp), -1.0, 1.0) == 0
        out[s] = row
    mkl.vslDeleteStream(ctypes.byref(st))
    return out


def write_mkl():
    mkl = _mkl()
    arrays = {}
    cases = []
    # (seed, n_imp, L, sample): the config seeds/shapes, small samples kept short by n
    for seed, n_imp, L, sample in [(1, 256, 4.0, 0), (1, 256, 4.0, 1), (7, 4096, 16.0, 0), (7, 4096, 16.0, 3),
                                   (3, 48000, 54.772255750516614, 1), (11, 1024, 8.0, 2), (0, 64, 2.0, 0),
                                   (4294967295, 64, 2.0, 5)]:
        n = min(n_imp, 512)              # keep the first n of each of the three draws
        skip = n_imp * 3 * sample
        x, y, b = rlm_stream(mkl, seed, skip, n_imp, L)
        key = f"rlm_s{seed}_n{n_imp}_k{sample}"
        arrays[key + "_x"] = x[:n]
        arrays[key + "_y"] = y[:n]
        arrays[key + "_bits"] = b[:n]
        cases.append((key, seed, n_imp, L, sample, n))
    for seed, n_sites, n_samples in [(1, 576, 3), (5, 9, 4)]:
        arrays[f"hlm_s{seed}_n{n_sites}"] = hlm_stream(mkl, seed, n_sites, n_samples)
    # FP32 builds: float L = sqrtf(n_phi * q), vsRngUniform (first 2048 of each draw kept)
    f32_cases = []
    for seed, n_phi, n_imp, sample in [(1, 16, 256, 1), (7, 256, 4096, 3), (3, 3000, 48000, 1), (11, 300, 4800, 2)]:
        L = float(np.sqrt(np.float32(n_phi)))
        x, y = rlm_stream_f32(mkl, seed, 3 * n_imp * sample, n_imp, L)
        key = f"rlm32_s{seed}_n{n_imp}_k{sample}"
        arrays[key + "_x"] = x[:2048]
        arrays[key + "_y"] = y[:2048]
        f32_cases.append((key, seed, n_phi, n_imp, sample, min(n_imp, 2048)))
    meta = {"mkl_version": _mkl_version(mkl), "rlm_cases": cases, "rlm_f32_cases": f32_cases}
    arrays["meta_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "mkl_rng.npz"), **arrays)
    print("wrote mkl_rng.npz:", meta["mkl_version"])


def write_oracle():
    from oracle import oracle as O
    out = {}
    # rlm C1 shape: -n 16 -q 1 -i 16 -g 64 -c 2 -s 1
    cases = [("rlm_n16_g64_s1", dict(n_phi=16, q=1.0, imp=16.0, n_mesh=64, seed=1), [0, 1]),
             ("rlm_n64_g16_s11", dict(n_phi=64, q=1.0, imp=16.0, n_mesh=16, seed=11), [0, 1]),
             ("rlm_n128_g16_s11", dict(n_phi=128, q=1.0, imp=16.0, n_mesh=16, seed=11), [0]),
             ("rlm_n256_g32_s7", dict(n_phi=256, q=1.0, imp=16.0, n_mesh=32, seed=7), [0, 1])]
    for name, kw, samples in cases:
        rows = []
        for s in samples:
            r = O.rlm_sample(kw["n_phi"], kw["q"], kw["imp"], kw["n_mesh"], kw["seed"], s, threads=8)
            rows.append({"sample": s, "energy": r.energy.tolist(), "chern": r.chern.tolist()})
            print(name, s, "sum C =", r.chern.sum(), flush=True)
        out[name] = {"params": kw, "samples": rows}
    # hlm C2 shape (24x24, q=8, b=0, g=32, seed 1), reference disorder width W=2
    lx = ly = 24
    Q, band, g, seed = 8, 0, 32, 1
    pots = O.hlm_disorder(seed, lx * ly, 2)
    rows = []
    for s in range(2):
        r = O.hlm_sample(lx, ly, Q, band, g, pots[s], threads=8)
        rows.append({"sample": s, "energy": r.energy.tolist(), "chern": r.chern.tolist()})
        print("hlm", s, "sum C =", r.chern.sum(), flush=True)
    out["hlm_24x24_q8_b0_g32_s1"] = {"params": dict(lx=lx, ly=ly, Q=Q, band=band, n_mesh=g, seed=seed),
                                     "samples": rows}
    with open(os.path.join(HERE, "oracle_fixtures.json"), "w") as f:
        json.dump(out, f)
    print("wrote oracle_fixtures.json")


def write_oracle_large():
    """Oracle outputs at the full configuration shapes that the GPU tests compare against
    without rerunning the CPU path: rlm n_phi=3000 g=4 seed 3 (samples 0-1), the uneven
    batch's n_phi = 192/256/384/512 at g=16 seed 11 (sample 0), and the hlm 24x24 q=8 b=0
    g=32 seed 1 configuration at disorder width W=1 (half of the reference's W=2 stream,
    samples 0-3). Each sample row also carries max |C - round C| of the oracle."""
    from oracle import oracle as O
    path = os.path.join(HERE, "oracle_fixtures_large.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)

    def save():
        with open(path, "w") as f:
            json.dump(out, f)

    def row(s, r):
        c = np.asarray(r.chern)
        return {"sample": s, "energy": r.energy.tolist(), "chern": c.tolist(),
                "max_int_dev": float(np.abs(c - np.rint(c)).max())}

    rlm_cases = [("rlm_n192_g16_s11", dict(n_phi=192, q=1.0, imp=16.0, n_mesh=16, seed=11), [0]),
                 ("rlm_n256_g16_s11", dict(n_phi=256, q=1.0, imp=16.0, n_mesh=16, seed=11), [0]),
                 ("rlm_n384_g16_s11", dict(n_phi=384, q=1.0, imp=16.0, n_mesh=16, seed=11), [0]),
                 ("rlm_n512_g16_s11", dict(n_phi=512, q=1.0, imp=16.0, n_mesh=16, seed=11), [0]),
                 ("rlm_n3000_g4_s3", dict(n_phi=3000, q=1.0, imp=16.0, n_mesh=4, seed=3), [0, 1])]
    for name, kw, samples in rlm_cases:
        have = {r["sample"] for r in out.get(name, {}).get("samples", [])}
        for s in samples:
            if s in have:
                continue
            r = O.rlm_sample(kw["n_phi"], kw["q"], kw["imp"], kw["n_mesh"], kw["seed"], s, threads=8)
            out.setdefault(name, {"params": kw, "samples": []})["samples"].append(row(s, r))
            print(name, s, "sum C =", r.chern.sum(), flush=True)
            save()
    name = "hlm_24x24_q8_b0_g32_s1_w1"
    lx = ly = 24
    Q, band, g, seed, S = 8, 0, 32, 1, 4
    pots = 0.5 * O.hlm_disorder(seed, lx * ly, S)       # W=1: half of the W=2 stream, exactly
    have = {r["sample"] for r in out.get(name, {}).get("samples", [])}
    for s in range(S):
        if s in have:
            continue
        r = O.hlm_sample(lx, ly, Q, band, g, pots[s], threads=8)
        out.setdefault(name, {"params": dict(lx=lx, ly=ly, Q=Q, band=band, n_mesh=g, seed=seed, width=1.0),
                              "samples": []})["samples"].append(row(s, r))
        print(name, s, "sum C =", r.chern.sum(), flush=True)
        save()
    print("wrote oracle_fixtures_large.json")


if __name__ == "__main__":
    what = sys.argv[1:] or ["mkl", "oracle"]
    if "mkl" in what:
        write_mkl()
    if "oracle" in what:
        write_oracle()
    if "large" in what:
        write_oracle_large()
