"""The drop-in drivers bin/rlm and bin/hlm (reference flags, energy_Chern_No.txt rows,
log files) against the CPU oracle: same row format as rlm/main_rlm_cpu.cpp:242 and
hlm/main_hlm_cpu.cpp:195, Chern integers identical, energies within 1e-10 relative
(compared on the printed 8 significant digits)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "random-landau-model_amd", "bin")


def _rows(path):
    out = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            out.append((int(parts[0]), float(parts[1]), float(parts[2]), line))
    return out


@pytest.mark.parametrize("prec", ["0", "1"])
def test_rlm_cli(tmp_path, oracle, prec):
    print(f"[test_rlm_cli] precision {prec}", flush=True)
    r = subprocess.run([os.path.join(BIN, "rlm"), "-n", "16", "-g", "8", "-c", "2", "-s", "1", "-t", "4", "-p", prec],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "estimated # of samples/hour:" in r.stderr
    log = (tmp_path / "log_rlm.txt").read_text()
    assert "seed: 1" in log and "n_phi: 16" in log
    rows = _rows(tmp_path / "energy_Chern_No.txt")
    assert len(rows) == 32
    for s in range(2):
        ref = oracle.rlm_sample(16, 1.0, 16.0, 8, 1, s)
        got = [x for x in rows if x[0] == s]
        e = np.array([x[1] for x in got])
        c = np.array([x[2] for x in got])
        assert np.abs(e - ref.energy).max() <= 1e-7 * np.abs(ref.energy).max()   # 8 printed digits
        assert np.array_equal(np.rint(c), np.rint(ref.chern))
        # the energy column is the reference's left-aligned setw(15) %.8g field: the FP64
        # path prints the oracle's text exactly; the mixed one (energies within 1e-12 of FP64,
        # which can move the 8th digit) prints the same layout
        for (_, _, _, line), ev in zip(got, ref.energy):
            field = line[9:24]
            assert line[:9] == f"{s:<8} " and line[24] == " ", line
            assert field == f"{'%.8g' % float(field):<15}", line
            if prec == "0":
                assert line.startswith(f"{s:<8} {'%.8g' % ev:<15} "), line


def test_rlm_cli_resume_from_sample(tmp_path):
    """-f resumes at a sample offset (rlm/disorder_potential.cpp:6 makes every sample
    addressable): samples 0-1 of one run then sample 2 of a resumed run equal samples 0-2
    of a single run, rows appended."""
    args = [os.path.join(BIN, "rlm"), "-n", "16", "-g", "6", "-s", "9"]
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    assert subprocess.run(args + ["-c", "3"], cwd=a, capture_output=True, timeout=120).returncode == 0
    assert subprocess.run(args + ["-c", "2"], cwd=b, capture_output=True, timeout=120).returncode == 0
    assert subprocess.run(args + ["-c", "1", "-f", "2"], cwd=b, capture_output=True, timeout=120).returncode == 0
    assert (a / "energy_Chern_No.txt").read_text() == (b / "energy_Chern_No.txt").read_text()


def test_rlm_cli_rejects_bad_flags(tmp_path):
    r = subprocess.run([os.path.join(BIN, "rlm"), "-z"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "Usage:" in r.stdout


def test_hlm_cli(tmp_path, oracle):
    print("[test_hlm_cli]", flush=True)
    r = subprocess.run([os.path.join(BIN, "hlm"), "-x", "6", "-y", "6", "-q", "3", "-g", "6", "-b", "0", "-c", "2",
                        "-s", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = _rows(tmp_path / "energy_Chern_No.txt")
    assert len(rows) == 24
    pots = oracle.hlm_disorder(1, 36, 2)
    for s in range(2):
        ref = oracle.hlm_sample(6, 6, 3, 0, 6, pots[s])
        got = [x for x in rows if x[0] == s]
        e = np.array([x[1] for x in got])
        c = np.array([x[2] for x in got])
        assert np.abs(e - ref.energy).max() <= 1e-7 * np.abs(ref.energy).max()
        assert np.array_equal(np.rint(c), np.rint(ref.chern))
        assert got[0][3].startswith(f"{s:>8} ")
    for line in ("diagonalization time/k-point:", "Chern No. calculation time/k-point:", "estimated total time:",
                 "estimated # of samples/hour:", "# module", "diagonalization", "Chern No. calculation"):
        assert line in r.stderr, line
