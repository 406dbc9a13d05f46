"""Python plumbing over librlm_hip.so (the C ABI declared in include/hip_diag.h).

Tests and bench.py drive the HIP library through these ctypes bindings; torch is
used only to own device memory. The library must be built
(``make -C random-landau-model_amd``); nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "librlm_hip.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "hip_diag.h")

HIP_DIAG_FP64 = 0
HIP_DIAG_MIXED = 1
HIP_DIAG_CONV_GPU_FP32 = 1

_lib = None


class HipDiagError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != 0:
        raise HipDiagError(f"{what} failed with status {rc}")


def lib() -> ctypes.CDLL:
    """Load librlm_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipDiagError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}`")
    # torch ships its own HIP runtime under the same soname as /opt/rocm's: load torch
    # first so the library binds to that copy (loading ours first leaves torch's HIP
    # initialisation without devices when the two runtimes differ)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.c_void_p
    i, l, d, vp = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
    sig = {
        "hip_diag_hamil": ([dp, dp, i], i),
        "hip_diag_hamil_v": ([dp, dp, i], None),
        "hip_diag_hamil_noalloc": ([dp, dp, i], i),
        "hip_heevd": ([dp, dp, i], i),
        "hip_diag_heevd_batched": ([i, dp, l, l, dp, l, i, ctypes.c_char, i, vp], i),
        "hip_cal_chern": ([dp, dp, i, i, i, i], i),
        "hip_write_wfs": ([ctypes.c_char_p, i, i, dp, ctypes.c_ulong, i], i),
        "hip_cal_chern_wfs_io": ([dp, i, i, i, i, ctypes.c_char_p, i], i),
        "rlm_plan_write_wfs": ([vp, i, i, ctypes.c_char_p, i], i),
        "rlm_prepare_coeff": ([dp, dp, dp, i, i, i, i, i, d, d], i),
        "rlm_generate_disorder_potential": ([dp, dp, dp, l, d, d, ctypes.c_ulong, l], i),
        "hip_generate_disorder_potential": ([dp, dp, dp, l, d, d, ctypes.c_ulong, l, i], i),
        "hip_hlm_generate_disorder": ([dp, l, l, ctypes.c_ulong, l, d], i),
        "hip_prepare_potential_coeff": ([dp, dp, dp, dp, dp, i, i, i, l], i),
        "hip_set_hamil_matrix": ([dp, dp, dp, dp, dp, i, i, i], i),
        "rlm_plan_create": ([ctypes.POINTER(vp), l, d, d, l, l, ctypes.c_ulong, i, i], i),
        "rlm_plan_create_ex": ([ctypes.POINTER(vp), l, d, d, l, l, ctypes.c_ulong, i, i, i], i),
        "rlm_plan_load": ([vp, l, i], i),
        "rlm_plan_execute": ([vp], i),
        "rlm_plan_fetch": ([vp, dp, dp, dp], i),
        "rlm_plan_run": ([vp, l, l, dp, dp, dp], i),
        "rlm_plan_stream": ([vp], vp),
        "rlm_plan_stage_times": ([vp, dp], i),
        "rlm_plan_info": ([vp, ctypes.POINTER(ctypes.c_long)], i),
        "rlm_plan_eigenvalues": ([vp], vp),
        "rlm_plan_max_residual": ([vp, dp], i),
        "rlm_plan_eigenvectors": ([vp], vp),
        "rlm_plan_refine_info": ([vp, ctypes.POINTER(ctypes.c_long)], i),
        "rlm_plan_pair_info": ([vp, ctypes.POINTER(ctypes.c_long)], i),
        "rlm_plan_loaded": ([vp], i),
        "rlm_plan_set_graph": ([vp, i], i),
        "rlm_plan_elapsed_ms": ([vp, dp], i),
        "rlm_multi_create": ([ctypes.POINTER(vp), i, ctypes.POINTER(ctypes.c_long), d, d, l, ctypes.c_ulong, i, i], i),
        "rlm_multi_load": ([vp, l, i], i),
        "rlm_multi_execute": ([vp], i),
        "rlm_multi_fetch": ([vp, i, dp, dp, dp], i),
        "rlm_multi_elapsed_ms": ([vp, dp], i),
        "rlm_multi_destroy": ([vp], i),
        "hip_diag_tuned_kernel": ([i, i], ctypes.c_char_p),
        "hip_diag_symv_timing": ([i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)], i),
        "rlm_plan_destroy": ([vp], i),
        "hlm_plan_create": ([ctypes.POINTER(vp), i, i, i, i, i, ctypes.c_ulong, d, i, i], i),
        "hlm_plan_load": ([vp, l, i], i),
        "hlm_plan_execute": ([vp], i),
        "hlm_plan_fetch": ([vp, dp, dp, dp], i),
        "hlm_plan_run": ([vp, l, l, dp, dp, dp], i),
        "hlm_plan_stage_times": ([vp, dp], i),
        "hlm_plan_stream": ([vp], vp),
        "hlm_plan_loaded": ([vp], i),
        "hlm_plan_destroy": ([vp], i),
        "hlm_generate_disorder": ([dp, l, l, ctypes.c_ulong, l, d], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def header_symbols() -> list[str]:
    """Function names declared in include/hip_diag.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- host helpers
def prepare_coeff(n_phi: int, off: int, d: int, n_mesh: int, L1: float, L2: float):
    """rlm_prepare_coeff -> (coeff_mn[d,d], coeff_m_theta[g+1,d], coeff_jm[N,d]) complex128."""
    cmn = np.zeros((d, d), np.complex128)
    cmt = np.zeros((n_mesh + 1, d), np.complex128)
    cjm = np.zeros((n_phi, d), np.complex128)
    _check(lib().rlm_prepare_coeff(_ptr(cmn), _ptr(cmt), _ptr(cjm), n_phi, off, d, d, n_mesh, L1, L2),
           "rlm_prepare_coeff")
    return cmn, cmt, cjm


def generate_disorder(n_imp: int, L1: float, L2: float, seed: int, sample: int):
    x = np.zeros(n_imp)
    y = np.zeros(n_imp)
    s = np.zeros(n_imp)
    _check(lib().rlm_generate_disorder_potential(_ptr(x), _ptr(y), _ptr(s), n_imp, L1, L2, seed, sample),
           "rlm_generate_disorder_potential")
    return x, y, s


def cal_chern(ring: np.ndarray, n_phi: int, dim_vec: int, n_mesh: int, theta_1: int) -> np.ndarray:
    """hip_cal_chern on a host ring [2, n_mesh+1, n_phi, dim_vec] (complex128)."""
    ring = np.ascontiguousarray(ring, dtype=np.complex128)
    out = np.zeros(n_phi * n_mesh)
    _check(lib().hip_cal_chern(_ptr(ring), _ptr(out), n_phi, dim_vec, n_mesh, theta_1), "hip_cal_chern")
    return out


def cal_chern_wfs_io(n_phi: int, dim_vec: int, n_mesh: int, theta_1: int, directory: str,
                     elem_bytes: int = 16) -> np.ndarray:
    """hip_cal_chern_wfs_io: the line's chern_theta from <dir>/wfs_<slot>_<theta_2>.bin."""
    out = np.zeros(n_phi * n_mesh)
    _check(lib().hip_cal_chern_wfs_io(_ptr(out), n_phi, dim_vec, n_mesh, theta_1, directory.encode(), elem_bytes),
           "hip_cal_chern_wfs_io")
    return out


# ---------------------------------------------------------------- device helpers (torch-owned memory)
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise HipDiagError("no GPU visible to torch")
    return torch


def heevd_batched(H, precision: int = HIP_DIAG_FP64, uplo: str = "L"):
    """Eigen-decomposition of a stack of Hermitian matrices H[b] (torch complex128, cuda).

    Returns (w[b, n], U[b, n, n]) with U[b][:, j] the eigenvector of w[b, j] (H U = U diag(w)).
    """
    torch = _torch()
    squeeze = H.dim() == 2
    if squeeze:
        H = H[None]
    b, n, _ = H.shape
    # the C ABI is column-major: hand it H^T's row-major buffer
    A = H.transpose(-1, -2).contiguous().to(torch.complex128)
    W = torch.empty((b, n), dtype=torch.float64, device=A.device)
    torch.cuda.synchronize()
    _check(lib().hip_diag_heevd_batched(n, A.data_ptr(), n, n * n, W.data_ptr(), n, b, uplo.encode(), precision,
                                        None), "hip_diag_heevd_batched")
    U = A.transpose(-1, -2)
    if squeeze:
        return W[0], U[0]
    return W, U


class RlmPlan:
    """rlm_plan_* (batched rlm pipeline)."""

    def __init__(self, n_phi, q=1.0, imp=16.0, n_mesh=32, seed=1, samples_per_batch=1, off_override=-1,
                 precision=HIP_DIAG_FP64, flags=0):
        L = lib()
        self._h = ctypes.c_void_p()
        _check(L.rlm_plan_create_ex(ctypes.byref(self._h), n_phi, q, imp, n_mesh, off_override, seed,
                                    samples_per_batch, precision, flags), "rlm_plan_create_ex")
        info = (ctypes.c_long * 6)()
        _check(L.rlm_plan_info(self._h, info), "rlm_plan_info")
        self.n_phi, self.n_imp, self.off, self.d, self.n_mesh, self.spb = list(info)

    @property
    def loaded(self) -> int:
        """Samples of the plan's current batch (the library's count, which sizes fetch())."""
        return int(lib().rlm_plan_loaded(self._h))

    def load(self, first_sample: int, count: int):
        _check(lib().rlm_plan_load(self._h, first_sample, count), "rlm_plan_load")

    def execute(self):
        _check(lib().rlm_plan_execute(self._h), "rlm_plan_execute")

    def fetch(self):
        n = self.loaded * self.n_phi
        e, c, m = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().rlm_plan_fetch(self._h, _ptr(e), _ptr(c), _ptr(m)), "rlm_plan_fetch")
        shp = (self.loaded, self.n_phi)
        return e.reshape(shp), c.reshape(shp), m.reshape(shp)

    def run(self, first_sample: int, n_samples: int):
        n = n_samples * self.n_phi
        e, c, m = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().rlm_plan_run(self._h, first_sample, n_samples, _ptr(e), _ptr(c), _ptr(m)), "rlm_plan_run")
        shp = (n_samples, self.n_phi)
        return e.reshape(shp), c.reshape(shp), m.reshape(shp)

    STAGES = ("potential+tables", "assembly", "tridiagonalisation", "divide_conquer", "backtransform", "refinement",
              "links", "plaquettes")

    def elapsed_ms(self) -> float:
        """rlm_plan_elapsed_ms: GPU time of the last execute."""
        out = np.zeros(1)
        _check(lib().rlm_plan_elapsed_ms(self._h, _ptr(out)), "rlm_plan_elapsed_ms")
        return float(out[0])

    def stage_times(self) -> dict:
        """Per-stage GPU times of the last execute ({} after an execute in graph mode)."""
        ms = np.zeros(len(self.STAGES))
        rc = lib().rlm_plan_stage_times(self._h, _ptr(ms))
        if rc == 0:
            return {}
        if rc != len(self.STAGES):
            raise HipDiagError(f"rlm_plan_stage_times failed with status {rc}")
        return dict(zip(self.STAGES, ms.tolist()))

    @property
    def stream(self) -> int:
        return lib().rlm_plan_stream(self._h) or 0

    def max_residual(self) -> float:
        out = np.zeros(1)
        _check(lib().rlm_plan_max_residual(self._h, _ptr(out)), "rlm_plan_max_residual")
        return float(out[0])

    def refine_info(self) -> tuple[bool, int]:
        """(refines, chunks cut or unfactorised in the last execute)."""
        out = (ctypes.c_long * 2)()
        _check(lib().rlm_plan_refine_info(self._h, out), "rlm_plan_refine_info")
        return bool(out[0]), int(out[1])

    def set_graph(self, enable: bool = True):
        """rlm_plan_set_graph: replay the execute's launches as one captured hipGraph."""
        _check(lib().rlm_plan_set_graph(self._h, 1 if enable else 0), "rlm_plan_set_graph")

    def write_wfs(self, sample: int, theta_1: int, directory: str, elem_bytes: int = 16):
        """rlm_plan_write_wfs: line theta_1 of batch sample `sample` -> <dir>/wfs_<theta_1%2>_<theta_2>.bin."""
        _check(lib().rlm_plan_write_wfs(self._h, sample, theta_1, directory.encode(), elem_bytes),
               "rlm_plan_write_wfs")

    def pair_info(self) -> dict:
        """rlm_plan_pair_info of the last execute."""
        out = (ctypes.c_long * 4)()
        _check(lib().rlm_plan_pair_info(self._h, out), "rlm_plan_pair_info")
        return dict(zip(("solved", "split", "dropped", "unrotated"), list(out)))

    def eigenvalues_ptr(self) -> int:
        return lib().rlm_plan_eigenvalues(self._h) or 0

    def eigenvectors_ptr(self) -> int:
        return lib().rlm_plan_eigenvectors(self._h) or 0

    def close(self):
        if self._h:
            lib().rlm_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuned_kernel(n: int, precision: int = HIP_DIAG_FP64) -> str:
    """hip_diag_tuned_kernel: the tridiagonalisation kernel the autotuner chose."""
    return lib().hip_diag_tuned_kernel(n, precision).decode()


def symv_timing(enable: bool):
    """hip_diag_symv_timing: start (enable=True, returns None) or stop and read
    (enable=False, returns (total_ms, launches)) the launch timing of the n > 512
    solver's Hermitian product kernel."""
    tot = ctypes.c_double(0.0)
    cnt = ctypes.c_long(0)
    _check(lib().hip_diag_symv_timing(1 if enable else 0, ctypes.byref(tot), ctypes.byref(cnt)),
           "hip_diag_symv_timing")
    return None if enable else (tot.value, cnt.value)


class RlmMulti:
    """rlm_multi_*: several n_phi in one hipGraph launch (one branch per size)."""

    def __init__(self, sizes, q=1.0, imp=16.0, n_mesh=16, seed=1, samples_per_batch=1, precision=HIP_DIAG_FP64):
        self._h = ctypes.c_void_p()
        self.sizes = [int(n) for n in sizes]
        arr = (ctypes.c_long * len(self.sizes))(*self.sizes)
        _check(lib().rlm_multi_create(ctypes.byref(self._h), len(self.sizes), arr, q, imp, n_mesh, seed,
                                      samples_per_batch, precision), "rlm_multi_create")
        self.loaded = 0

    def load(self, first_sample: int, count: int):
        # a load that fails part-way leaves no count behind, so fetch() refuses rather than
        # sizing its buffers from a stale one
        self.loaded = 0
        _check(lib().rlm_multi_load(self._h, first_sample, count), "rlm_multi_load")
        self.loaded = count

    def execute(self):
        _check(lib().rlm_multi_execute(self._h), "rlm_multi_execute")

    def fetch(self, k: int):
        if self.loaded < 1:
            raise HipDiagError("rlm_multi_fetch: no batch loaded")
        n = self.loaded * self.sizes[k]
        e, c, m = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().rlm_multi_fetch(self._h, k, _ptr(e), _ptr(c), _ptr(m)), "rlm_multi_fetch")
        shp = (self.loaded, self.sizes[k])
        return e.reshape(shp), c.reshape(shp), m.reshape(shp)

    def elapsed_ms(self) -> float:
        out = np.zeros(1)
        _check(lib().rlm_multi_elapsed_ms(self._h, _ptr(out)), "rlm_multi_elapsed_ms")
        return float(out[0])

    def close(self):
        if self._h:
            lib().rlm_multi_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hlm_disorder(n_sites: int, n_samples: int, seed: int, first_sample: int = 0, W: float = 2.0):
    V = np.zeros(n_sites * n_samples)
    _check(lib().hlm_generate_disorder(_ptr(V), n_sites, n_samples, seed, first_sample, W), "hlm_generate_disorder")
    return V.reshape(n_samples, n_sites)


def device_disorder(n_imp: int, L1: float, L2: float, seed: int, first_sample: int, n_samples: int):
    """hip_generate_disorder_potential -> numpy (x, y, s), each [n_samples, n_imp]."""
    torch = _torch()
    out = [torch.zeros(max(n_samples * n_imp, 1), dtype=torch.float64, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    _check(lib().hip_generate_disorder_potential(out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), n_imp, L1, L2,
                                                 seed, first_sample, n_samples), "hip_generate_disorder_potential")
    return tuple(t[:n_samples * n_imp].cpu().numpy().reshape(n_samples, n_imp) for t in out)


def device_hlm_disorder(n_sites: int, n_samples: int, seed: int, first_sample: int = 0, W: float = 2.0):
    """hip_hlm_generate_disorder -> numpy [n_samples, n_sites]."""
    torch = _torch()
    V = torch.zeros(max(n_samples * n_sites, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _check(lib().hip_hlm_generate_disorder(V.data_ptr(), n_sites, n_samples, seed, first_sample, W),
           "hip_hlm_generate_disorder")
    return V[:n_samples * n_sites].cpu().numpy().reshape(n_samples, n_sites)


class HlmPlan:
    """hlm_plan_* (batched hlm pipeline)."""

    STAGES = ("projection", "tridiagonalisation", "divide_conquer", "backtransform", "links", "plaquettes", "unused")

    def __init__(self, lx, ly, Q, band=0, n_mesh=32, seed=1, W=2.0, samples_per_batch=1, precision=HIP_DIAG_FP64):
        self._h = ctypes.c_void_p()
        _check(lib().hlm_plan_create(ctypes.byref(self._h), lx, ly, Q, band, n_mesh, seed, W, samples_per_batch,
                                     precision), "hlm_plan_create")
        self.n_phi = lx * ly // Q
        self.spb = samples_per_batch

    @property
    def loaded(self) -> int:
        """Samples of the plan's current batch (the library's count, which sizes fetch())."""
        return int(lib().hlm_plan_loaded(self._h))

    def load(self, first_sample: int, count: int):
        _check(lib().hlm_plan_load(self._h, first_sample, count), "hlm_plan_load")

    def execute(self):
        _check(lib().hlm_plan_execute(self._h), "hlm_plan_execute")

    def fetch(self):
        n = self.loaded * self.n_phi
        e, c, m = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().hlm_plan_fetch(self._h, _ptr(e), _ptr(c), _ptr(m)), "hlm_plan_fetch")
        shp = (self.loaded, self.n_phi)
        return e.reshape(shp), c.reshape(shp), m.reshape(shp)

    def run(self, first_sample: int, n_samples: int):
        n = n_samples * self.n_phi
        e, c, m = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().hlm_plan_run(self._h, first_sample, n_samples, _ptr(e), _ptr(c), _ptr(m)), "hlm_plan_run")
        shp = (n_samples, self.n_phi)
        return e.reshape(shp), c.reshape(shp), m.reshape(shp)

    def stage_times(self) -> dict:
        ms = np.zeros(len(self.STAGES))
        rc = lib().hlm_plan_stage_times(self._h, _ptr(ms))
        if rc != len(self.STAGES):
            raise HipDiagError(f"hlm_plan_stage_times failed with status {rc}")
        return dict(zip(self.STAGES, ms.tolist()))

    def close(self):
        if self._h:
            lib().hlm_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
