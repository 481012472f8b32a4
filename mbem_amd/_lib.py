"""Loader of libmbem_hip.so (the C ABI declared in include/mbem_hip.h).

torch is imported first so that the HIP runtime, rocBLAS and rocSOLVER the
library links against resolve to the copies torch already loaded (same
SONAMEs): one runtime per process, so torch-allocated device memory and
torch streams are valid arguments.  There is no CPU fallback: a missing
library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch  # noqa: F401  (must precede the library load, see docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbem_hip.so")

MBEM_OK, MBEM_EINVAL, MBEM_EHIP, MBEM_ESCRATCH = 0, 1, 2, 3
NOALLOC_SCRATCH_DOUBLES = 1 << 20      # MBEM_NOALLOC_SCRATCH_DOUBLES
BJ_ROWS_PER_WAVE = 4
BJ32_ROWS_PER_WAVE = 8

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_f = ctypes.c_float
_int = ctypes.c_int

# name -> argument types (every entry returns int status unless listed below)
_SIGS = {
    "mbem_prepare_sources": [_p, _p, _p, _i64, _p, _i64, _p],
    "mbem_u_block": [_p, _i64, _p, _i64, _i64, _p, _p, _i64, _d, _d, _p],
    "mbem_t_block": [_p, _i64, _p, _i64, _i64, _p, _p, _i64, _d, _d, _p, _d, _int, _p],
    "mbem_u_basis": [_p, _i64, _p, _i64, _i64, _p, _i64, _i64, _p],
    "mbem_t_basis": [_p, _i64, _p, _i64, _i64, _p, _i64, _i64, _p],
    "mbem_u_disp_contract": [_p, _i64, _p, _i64, _i64, _p, _p, _p, _d, _d, _p],
    "mbem_t_disp_contract": [_p, _i64, _p, _i64, _i64, _p, _p, _p, _d, _d, _p],
    "mbem_kelvin_stress_contract": [_p, _i64, _p, _i64, _i64, _p, _d, _d, _p, _d, _d, _p],
    "mbem_dd_stress_contract": [_p, _i64, _p, _i64, _i64, _p, _d, _d, _p, _d, _d, _p],
    "mbem_eigenstress_contract": [_p, _i64, _p, _i64, _i64, _p, _d, _d, _p, _d, _d, _p],
    "mbem_eigenstress_near": [_p, _i64, _p, _i64, _i64, _p, _p, _p, _p, _p, _d, _d, _p, _d, _d, _p],
    "mbem_add_block_diag": [_p, _i64, _i64, _p, _d, _p],
    "mbem_gemv": [_p, _i64, _i64, _i64, _p, _p, _d, _d, _p],
    "mbem_multidot": [_p, _i64, _i64, _i64, _p, _p, _p, _p],
    "mbem_multiaxpy": [_p, _i64, _i64, _i64, _p, _p, _d, _p],
    "mbem_multidot_noalloc": [_p, _i64, _i64, _i64, _p, _p, _p],
    "mbem_axpby": [_i64, _d, _p, _d, _p, _p],
    "mbem_normalize": [_i64, _p, _p, _p],
    "mbem_gather_block": [_p, _i64, _p, _i64, _p, _p],
    "mbem_invert": [_p, _i64, _p, _p, _p],
    "mbem_invert_batched": [_p, _i64, _i64, _p, _p, _p],
    "mbem_invert_batched_oop": [_p, _i64, _i64, _p, _p, _p, _int, _p],
    "mbem_gather_block_pad": [_p, _i64, _p, _i64, _i64, _p, _int, _p],
    "mbem_copy2d": [_p, _i64, _p, _i64, _i64, _i64, _int, _p],
    "mbem_lu_factor": [_p, _i64, _p, _p, _p],
    "mbem_lu_solve": [_p, _i64, _p, _p, _i64, _p],
    "mbem_lu_solve_adjoint": [_p, _i64, _p, _p, _i64, _p],
    "mbem_lu_inv_doubles": [_i64],
    "mbem_lu_prepare": [_p, _i64, _p, _p, _p],
    "mbem_lu_solve_blocked": [_p, _i64, _p, _p, _p, _p, _p, _p, _int, _p],
    "mbem_lu_factor_f32": [_p, _i64, _p, _p, _p],
    "mbem_lu_prepare_f32": [_p, _i64, _p, _p, _p],
    "mbem_lu_solve_blocked_f32": [_p, _i64, _p, _p, _p, _p, _p, _p, _int, _p],
    "mbem_colsum_work": [_i64],
    "mbem_colsum_abs": [_p, _i64, _i64, _i64, _p, _p, _p],
    "mbem_bj_apply": [_p, _i64, _p, _p, _p, _p, _p, _p, _p],
    "mbem_cast_lo": [_p, _i64, _i64, _i64, _p, _i64, _int, _p],
    "mbem_near_pattern": [_p, _p, _p, _i64, _p, _p, _p, _p],
    "mbem_near_extract": [_p, _i64, _i64, _p, _p, _i64, _p, _p, _i64, _int, _p],
    "mbem_vec_to_f32": [_p, _i64, _p, _p],
    "mbem_gemv_lo": [_p, _i64, _i64, _i64, _p, _p, _d, _d, _int, _p],
    "mbem_i8_segments": [_i64],
    "mbem_tri_geometry": [_p, _i64, _p, _i64, _p, _p, _p],
    "mbem_cast_i8": [_p, _i64, _i64, _i64, _p, _i64, _p, _i64, _p],
    "mbem_near_hold": [_p, _i64, _i64, _p, _p, _i64, _p, _int, _p],
    "mbem_gemv_i8": [_p, _i64, _p, _i64, _i64, _i64, _p, _p, _d, _d, _p],
    "mbem_i8e_resid_bytes": [_i64],
    "mbem_cast_i8e": [_p, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _p],
    "mbem_gemv_i8e": [_p, _i64, _p, _i64, _p, _i64, _i64, _i64, _p, _p, _d, _d, _p],
    "mbem_bsr_gemv": [_i64, _p, _p, _p, _p, _p, _d, _p],
    "mbem_gather_block_f32": [_p, _i64, _p, _i64, _p, _p],
    "mbem_invert_batched_f32": [_p, _i64, _i64, _p, _p, _p],
    "mbem_bj_apply_f32": [_p, _i64, _p, _p, _p, _p, _p, _p, _p],
    "mbem_gather_f32": [_p, _i64, _p, _p, _p],
    "mbem_bj_apply_f32v": [_p, _i64, _p, _p, _p, _p, _p, _p, _p, _p],
    "mbem_bj_desc_fields": [],
    "mbem_bj_panel_width": [],
    "mbem_bj_gather_f32": [_p, _i64, _p, _i64, _p, _p, _i64, _i64, _p, _i64, _p, _i64, _p, _p],
    "mbem_bj_invert_f32": [_p, _p, _i64, _i64, _p, _i64, _p, _p, _i64, _p, _i64, _p, _p, _i64,
                           _p],
    "mbem_src_fields": [],
    "mbem_multidot_work": [_i64, _i64],
    # batched Laplace-state solves
    "mbem_mgemm_configs": [],
    "mbem_mgemm_config": [_int, _p],
    "mbem_mgemm_config_fits": [_int, _int],
    "mbem_nodal_geo_doubles": [_i64],
    "mbem_nodal_table_doubles": [_i64, _i64, _i64],
    "mbem_nodal_frames": [_p, _i64, _p, _p],
    "mbem_nodal_matrix": [_int, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _i64, _i64, _p],
    "mbem_nodal_contract": [_int, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _d, _d, _d,
                            _p, _p],
    "mbem_nodal_traction": [_int, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _i64, _d, _d,
                            _d, _p, _i64, _p],
    "mbem_gemm_rm": [_p, _i64, _i64, _i64, _p, _i64, _i64, _p, _i64, _d, _d, _int, _p],
    "mbem_gemm_g": [_p, _p, _i64, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _d, _d, _int, _p],
    "mbem_work_item_bytes": [],
    "mbem_dgemm_rowmajor": [_i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _p],
    "mbem_dgemm_abt": [_i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _p],
    "mbem_gemm_grouped": [_p, _p, _i64, _p, _i64, _p, _i64, _int, _p],
    "mbem_bdot": [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p],
    "mbem_breduce": [_p, _i64, _i64, _i64, _p, _i64, _i64, _int, _p, _p],
    "mbem_baxpy": [_p, _i64, _i64, _p, _p, _i64, _d, _p, _i64, _i64, _p],
    "mbem_bscale": [_p, _p, _p, _i64, _i64, _p],
    "mbem_bcopy": [_p, _p, _i64, _p],
    "mbem_fill": [_p, _i64, _d, _p],
    "mbem_fg_start": [_i64, _i64, _p, _p, _p, _p, _p, _p],
    "mbem_fg_givens": [_i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "mbem_fg_solve": [_i64, _i64, _p, _p, _p, _p, _p],
    "mbem_gather_comb": [_p, _p, _i64, _p, _i64, _d, _p, _p],
    "mbem_gather_comb_pad": [_p, _p, _i64, _p, _i64, _i64, _d, _p, _p],
    "mbem_permute": [_p, _i64, _p, _i64, _p, _i64, _p],
    "mbem_graph_begin": [_p],
    "mbem_graph_end": [_p, _p],
    "mbem_graph_launch": [_p, _p],
    "mbem_graph_destroy": [_p],
    "mbem_memcpy_async": [_p, _p, _i64, _p],
    # coarse space of the two-level block-Jacobi
    "mbem_coarse_image": [_p, _i64, _i64, _p, _p, _p, _i64, _int, _p, _i64, _p],
    "mbem_coarse_expand": [_p, _p, _p, _p, _i64, _int, _d, _p, _p],
    "mbem_gram_work": [_i64, _i64],
    "mbem_coarse_project_work": [_i64, _i64],
    "mbem_coarse_project_f32": [_p, _i64, _i64, _i64, _p, _p, _p, _p, _p],
    "mbem_coarse_image_i8": [_p, _i64, _p, _i64, _i64, _p, _p, _p, _i64, _int, _p, _i64, _p],
    "mbem_coarse_image_bsr": [_i64, _p, _p, _p, _p, _p, _p, _int, _p, _i64, _p],
    "mbem_gram": [_p, _i64, _i64, _i64, _p, _p, _p],
    "mbem_coarse_factor": [_p, _i64, _d, _p, _p, _p, _p, _p],
    # learned preconditioner
    "mbem_knn": [_p, _i64, _int, _p, _p],
    "mbem_sgemm": [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64, _i64, _i64, _f, _f, _p, _int,
                   _i64, _p, _p],
    "mbem_colsum": [_p, _i64, _i64, _p, _p, _i64, _f, _p],
    "mbem_knn_mean": [_p, _i64, _i64, _p, _int, _p, _p],
    "mbem_knn_mean_bwd": [_p, _i64, _i64, _p, _p, _int, _p, _p],
    "mbem_relu_bwd": [_p, _p, _i64, _p],
    "mbem_concat": [_p, _i64, _p, _i64, _i64, _p, _p],
    "mbem_split": [_p, _i64, _i64, _i64, _p, _int, _p, _p],
    "mbem_add_row": [_p, _i64, _i64, _p, _f, _p],
    "mbem_loss": [_p, _p, _i64, _f, _p, _p, _p, _p],
    "mbem_adam": [_p, _p, _p, _p, _i64, _f, _f, _f, _f, _f, _p],
    "mbem_hgemm": [_p, _i64, _p, _i64, _i64, _i64, _p, _int, _p, _i64, _p],
    "mbem_lr_coef": [_p, _i64, _i64, _i64, _p, _i64, _i64, _int, _p, _i64, _p, _p],
    "mbem_lr_expand": [_p, _i64, _i64, _i64, _p, _i64, _p, _i64, _int, _d, _int, _p, _p],
    "mbem_lr_outer": [_p, _i64, _p, _i64, _i64, _i64, _i64, _p, _p],
    "mbem_lr_loss": [_p, _p, _i64, _i64, _p, _p, _p],
}
_RESTYPE = {"mbem_nodal_geo_doubles": _i64, "mbem_nodal_table_doubles": _i64,
            "mbem_multidot_work": _i64, "mbem_gram_work": _i64, "mbem_i8_segments": _i64, "mbem_colsum_work": _i64, "mbem_lu_inv_doubles": _i64,
            "mbem_i8e_resid_bytes": _i64, "mbem_coarse_project_work": _i64}

_lib = None


class MbemError(RuntimeError):
    pass


def exported_names() -> list:
    """Every C entry point this module binds (include/mbem_hip.h)."""
    return sorted(_SIGS)


def load():
    """The bound library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              f"`python -m mbem_amd.build` (hipcc, gfx950)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, _int)
        _lib = lib
    return _lib


def require_gpu():
    if not torch.cuda.is_available():
        raise MbemError("libmbem_hip needs a ROCm GPU (torch.cuda.is_available() is False); "
                        "mbem_amd has no CPU path")
    load()


def device():
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


class _StagingRing:
    """One pinned host buffer that every upload stages through, as a ring.

    Each upload writes its bytes at the ring's head (a plain memcpy), queues
    the copy to the device on the current stream and records an event after
    it; a region is written again only once its event has completed, so the
    host waits only if it laps a copy the device has not reached.  A fresh
    pinned tensor per upload instead goes through torch's pinned-block cache,
    whose blocks cannot be reused while their copies are still queued (the
    usual case here: uploads are queued behind the assembly), so it pins new
    memory per upload, at 0.3 ms each on MI355X and up to several ms now and
    then.  Uploads larger than half the ring take that path."""

    CAPACITY = 64 << 20
    ALIGN = 256

    def __init__(self):
        self.buf = torch.empty(self.CAPACITY, dtype=torch.uint8, pin_memory=True)
        self.host = self.buf.numpy()
        self.head = 0
        self.pending = []               # (start, end, event), oldest first

    def _claim(self, nbytes: int) -> int:
        start = self.head
        if start + nbytes > self.CAPACITY:
            start = 0
        end = start + nbytes
        keep = []
        for s, e, ev in self.pending:
            if s < end and start < e:
                ev.synchronize()        # the device has not copied this region yet
            elif not ev.query():
                keep.append((s, e, ev))
        self.pending = keep
        self.head = end + (-end % self.ALIGN)
        return start

    def upload(self, a: np.ndarray, dtype) -> torch.Tensor:
        npdt = _np_dtype(dtype)
        nbytes = a.size * npdt.itemsize
        dev = device()
        if nbytes == 0:
            return torch.empty(a.shape, dtype=dtype, device=dev)
        start = self._claim(nbytes)
        stage = self.host[start:start + nbytes].view(npdt).reshape(a.shape)
        np.copyto(stage, a, casting="unsafe")
        out = torch.empty(a.shape, dtype=dtype, device=dev)
        out.copy_(self.buf[start:start + nbytes].view(dtype).view(a.shape), non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self.pending.append((start, start + nbytes, ev))
        return out


_NP_DTYPES = {}
_ring = None


def _np_dtype(dtype) -> np.dtype:
    if dtype not in _NP_DTYPES:
        _NP_DTYPES[dtype] = torch.empty(0, dtype=dtype).numpy().dtype
    return _NP_DTYPES[dtype]


def upload(a, dtype):
    """A host array on the device through a pinned staging copy, without
    waiting for the work already queued on the stream (the staging ring's
    region is reused only after the device has copied it)."""
    global _ring
    a = np.ascontiguousarray(a)         # (a 0-d array becomes shape (1,), as torch saw it)
    if a.size * _np_dtype(dtype).itemsize > _StagingRing.CAPACITY // 2:
        t = torch.as_tensor(a, dtype=dtype).pin_memory()
        return t.to(device(), non_blocking=True)
    if _ring is None:
        _ring = _StagingRing()
    return _ring.upload(a, dtype)


_raw_stream = None


def stream():
    """hipStream_t of torch's current stream on the current device (the raw
    handle torch's own graph tracer reads; the first call goes through
    ``torch.cuda.current_stream`` so the runtime is initialised)."""
    global _raw_stream
    if _raw_stream is None:
        torch.cuda.current_stream()
        get, dev = torch._C._cuda_getCurrentRawStream, torch._C._cuda_getDevice
        _raw_stream = lambda: get(dev())  # noqa: E731
    return ctypes.c_void_p(_raw_stream())


def ptr(t) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def host_ptr(a: np.ndarray) -> ctypes.c_void_p:
    """Address of a C-contiguous host array (for the library's host functions)."""
    if not a.flags.c_contiguous:
        raise ValueError("host arrays passed to the library must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def check(status: int, what: str) -> None:
    if status == MBEM_OK:
        return
    kind = {MBEM_EINVAL: "invalid argument", MBEM_EHIP: "HIP error",
            MBEM_ESCRATCH: "needs more than the library's fixed device scratch"}.get(
        status, f"status {status}")
    raise MbemError(f"{what}: {kind}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def host_doubles(vals):
    arr = (ctypes.c_double * len(vals))(*[float(v) for v in vals])
    return arr


SRC_FIELDS = 24
