"""ctypes binding of libmr (include/mr.h) for tests, the benchmark and scripts.

Mirrors the reference's GRAPH<V> operator surface (include/class.cuh:59-99):
``Graph`` wraps a matrix handle with ``gorder6`` / ``gorder1`` / ``relabel_rcm``
/ ``locality_predict`` / ``spmv``; the lower-level ``Mat`` exposes every
C-ABI entry point.  Device memory is handled through the library's own
plumbing (``Ctx.malloc`` / ``DeviceArray``), so no other HIP runtime has to be
loaded into the process.  Loading fails loudly when libmr.so is missing: there
is no CPU fallback for any product entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import weakref

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent.parent
LIB_PATH = PKG / "lib" / "libmr.so"

F64, F32, F16, BF16 = 0, 1, 2, 3
MERBIT, CSR_SCALAR, CSR_VECTOR = 0, 1, 2
H2D, D2H, D2D = 0, 1, 2
RD_EXACT, RD_SAMPLED = 0, 1
MODEL_MERBIT, MODEL_CSR, MODEL_MERBIT_MI355X, MODEL_CSR_MI355X = 0, 1, 2, 3

_NP_OF = {F64: np.float64, F32: np.float32, F16: np.float16, BF16: np.uint16}
STATUS = {
    0: "MR_OK", 1: "MR_ERR_INVALID_ARG", 2: "MR_ERR_HIP", 3: "MR_ERR_OUT_OF_MEMORY",
    4: "MR_ERR_NOT_PERMUTATION", 5: "MR_ERR_UNSUPPORTED", 6: "MR_ERR_MALFORMED_CSR", 7: "MR_ERR_IO",
    8: "MR_ERR_ADMISSION", 9: "MR_ERR_NO_DEVICE", 10: "MR_ERR_NUMERIC", 11: "MR_ERR_TIMEOUT",
    12: "MR_ERR_CANCELLED",
}


class MRError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class ReorderStats(C.Structure):
    _fields_ = [("width", C.c_int32), ("window", C.c_int32), ("batches", C.c_int32),
                ("gpu_ms", C.c_double), ("cpu_ms", C.c_double), ("relabel_ms", C.c_double),
                ("total_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RDParams(C.Structure):
    _fields_ = [("sector_size", C.c_int32), ("samples", C.c_int32), ("rd_cap", C.c_int32),
                ("hll_p", C.c_int32), ("max_scan", C.c_int32), ("stream_type", C.c_int32),
                ("graph_id", C.c_int32), ("n_edges", C.c_int32), ("bin_edges", C.c_int32 * 32)]


class RDStats(C.Structure):
    _fields_ = [("reuses", C.c_int64), ("cold", C.c_int64), ("found", C.c_int32),
                ("cold_samples", C.c_int32), ("large", C.c_int32), ("unresolved", C.c_int32),
                ("ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int), ("max_jobs", C.c_int), ("coalesce", C.c_int), ("verify", C.c_int),
                ("sigma", C.c_int), ("grid", C.c_int), ("upload_threads", C.c_int), ("heavy_slots", C.c_int),
                ("slot_bytes", C.c_int64), ("hbm_budget", C.c_int64), ("hbm_reserve", C.c_int64),
                ("timeout_ms", C.c_double)]


class ServeLoad(C.Structure):
    _fields_ = [("rate", C.c_double), ("n_requests", C.c_int64), ("seed", C.c_uint64),
                ("k_choices", C.c_int * 4), ("n_k", C.c_int), ("upload_every", C.c_int),
                ("upload_scale", C.c_int), ("upload_edge_factor", C.c_int)]


class ServeReport(C.Structure):
    _fields_ = [("completed", C.c_int64), ("rejected", C.c_int64), ("timed_out", C.c_int64),
                ("cancelled", C.c_int64), ("uploads_admitted", C.c_int64), ("uploads_rejected", C.c_int64),
                ("evictions", C.c_int64), ("p50_us", C.c_double), ("p90_us", C.c_double), ("p99_us", C.c_double),
                ("mean_us", C.c_double), ("max_us", C.c_double), ("wall_s", C.c_double), ("gflops", C.c_double),
                ("gbytes_s", C.c_double), ("offered_rate", C.c_double), ("achieved_rate", C.c_double),
                ("batches", C.c_double), ("mean_jobs_per_batch", C.c_double),
                ("mean_requests_per_batch", C.c_double), ("upload_ms_mean", C.c_double),
                ("completed_heavy", C.c_int64), ("timed_out_heavy", C.c_int64), ("p99_light_us", C.c_double),
                ("p99_heavy_us", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)

# name -> argtypes (every entry point returns int unless listed in _RESTYPE)
SIGNATURES = {
    "mr_last_error": [], "mr_version": [],
    "mr_device_count": [C.POINTER(C.c_int)],
    "mr_ctx_create": [C.c_int, C.POINTER(_P)], "mr_ctx_destroy": [_P],
    "mr_ctx_stream": [_P, C.POINTER(_P)],
    "mr_ctx_mem_info": [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "mr_ctx_arch": [_P, C.c_char_p, C.c_int, C.POINTER(C.c_int)],
    "mr_malloc": [_P, C.c_size_t, C.POINTER(_P)], "mr_free": [_P, _P],
    "mr_memcpy": [_P, _P, _P, C.c_size_t, C.c_int], "mr_memset": [_P, _P, C.c_int, C.c_size_t],
    "mr_bench_copy": [_P, _P, _P, C.c_size_t, _P],
    "mr_stream_create": [_P, C.POINTER(_P)], "mr_stream_destroy": [_P, _P], "mr_stream_sync": [_P, _P],
    "mr_event_create": [_P, C.POINTER(_P)], "mr_event_destroy": [_P, _P],
    "mr_event_record": [_P, _P, _P], "mr_event_elapsed_ms": [_P, _P, _P, _FP],
    "mr_gen_rmat_edges": [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                          _I32P, _I32P],
    "mr_coo_to_csr": [C.c_int64, C.c_int64, _I32P, _I32P, C.c_int, _I32P, _I32P, _I64P],
    "mr_gen_values": [C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_int64, _DP],
    "mr_gen_stencil27": [C.c_int, C.c_int, C.c_int, _I32P, _I32P, _I64P, _I64P],
    "mr_gen_banded": [C.c_int64, C.c_int, C.c_int, C.c_uint64, _I32P, _I32P, _I64P],
    "mr_csr_create": [_P, C.c_int64, C.c_int64, C.c_int64, _I32P, _I32P, _P, C.c_int, C.c_int, C.POINTER(_P)],
    "mr_mtx_load": [_P, C.c_char_p, C.c_int, C.c_int, C.POINTER(_P)],
    "mr_gap_load": [_P, C.c_char_p, C.c_int, C.c_int, C.POINTER(_P)],
    "mr_mat_destroy": [_P],
    "mr_mat_info": [_P, _I64P, _I64P, _I64P, C.POINTER(C.c_int)],
    "mr_mat_export": [_P, _I32P, _I32P, _P],
    "mr_mat_convert": [_P, C.c_int, C.POINTER(_P)],
    "mr_mat_perm": [_P, _I32P],
    "mr_check_perm": [_I32P, C.c_int64],
    "mr_permute": [_P, _I32P],
    "mr_permute_vec": [_P, C.c_int, _P, _P, _P],
    "mr_get_width": [C.c_int64, C.c_int64, C.POINTER(C.c_int)],
    "mr_reorder_pgorder": [_P, C.c_int, C.c_int, _I32P, C.POINTER(ReorderStats)],
    "mr_pgorder_order": [_P, C.c_int, C.c_int, _U32P, _U32P, C.POINTER(ReorderStats)],
    "mr_pgorder_cpu_phase": [_P, C.c_int, C.c_int, _U32P, _U32P, C.c_int, _DP],
    "mr_reorder_gorder_cpu": [_P, C.c_int, _I32P, C.POINTER(ReorderStats)],
    "mr_reorder_rcm": [_P, _I32P, C.POINTER(ReorderStats)],
    "mr_reorder_rabbit": [_P, C.c_int, _I32P, C.POINTER(ReorderStats)],
    "mr_reorder_gorder5": [_P, C.c_int, C.c_int, C.c_int, _I32P, C.POINTER(ReorderStats)],
    "mr_gorder5_order": [_P, C.c_int, C.c_int, C.c_int, _U32P, _U32P, C.POINTER(ReorderStats)],
    "mr_reorder_tile_order": [_P, C.c_int, C.c_int, _I32P, C.POINTER(ReorderStats)],
    "mr_reorder_degree": [_P, _I32P, C.POINTER(ReorderStats)],
    "mr_reorder_dbg": [_P, _I32P, C.POINTER(ReorderStats)],
    "mr_csr_hash": [C.c_int64, C.c_int64, C.c_int64, _I32P, _I32P, C.POINTER(C.c_uint64)],
    "mr_mat_hash": [_P, C.POINTER(C.c_uint64)],
    "mr_reorder_pgorder_cached": [_P, C.c_int, C.c_int, C.c_char_p, _I32P, C.POINTER(ReorderStats),
                                  C.POINTER(C.c_int)],
    "mr_plan_save": [_P],
    "mr_rd_params_default": [C.POINTER(RDParams)],
    "mr_rd_params_model": [C.c_int, C.POINTER(RDParams)],
    "mr_rd_exact_stream": [_I32P, C.c_int64, C.c_int64, C.POINTER(RDParams), _I64P, _DP, C.POINTER(RDStats)],
    "mr_rd_features": [_P, C.c_int, C.POINTER(RDParams), _DP, C.POINTER(RDStats)],
    "mr_rd_sampled_raw": [_P, C.POINTER(RDParams), _FP, _I32P, _I32P],
    "mr_predict": [_DP, C.c_int, _DP],
    "mr_locality_predict": [_P, C.c_int, _DP, _DP, C.POINTER(RDStats)],
    "mr_spmv": [_P, C.c_int, _P, _P, _P],
    "mr_merbit_ws_bytes": [_P, C.POINTER(C.c_size_t)],
    "mr_spmv_ws": [_P, _P, _P, _P, _P],
    "mr_spmv_noalloc": [_P, _P, _P, _P],
    "mr_spmm": [_P, C.c_int, C.c_int, _P, C.c_int64, _P, C.c_int64, _P],
    "mr_merbit_configure": [_P, C.c_int],
    "mr_merbit_options": [_P, C.c_int],
    "mr_merbit_get_options": [_P, C.POINTER(C.c_int)],
    "mr_merbit_autotune": [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _FP],
    "mr_merbit_info": [_P, C.POINTER(C.c_int), _I64P, _I64P, _DP],
    "mr_merbit_export": [_P, _U32P, _I32P, _I32P],
    "mr_merbit_xc_configure": [_P, C.c_int, C.c_int, C.c_int],
    "mr_merbit_panel_configure": [_P, C.c_int, C.c_int],
    "mr_merbit_panel_info": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), _I64P, _I64P, _I64P, C.POINTER(C.c_int),
                             _DP],
    "mr_merbit_stored_bytes": [_P, _I64P],
    "mr_merbit_xc_info": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _I64P],
    "mr_batch_create": [_P, C.c_int, C.POINTER(_P), C.c_int, C.POINTER(_P)],
    "mr_batch_destroy": [_P],
    "mr_batch_set_io": [_P, C.POINTER(_P), _I64P, C.POINTER(_P), _I64P, _P],
    "mr_batch_configure": [_P, C.c_int, C.c_int, C.c_int],
    "mr_batch_autotune": [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _FP],
    "mr_batch_run": [_P, _P],
    "mr_batch_graph_capture": [_P, _P],
    "mr_batch_graph_launch": [_P, _P],
    "mr_batch_info": [_P, _I64P, _I64P, _DP, _DP],
    "mr_serve_config_default": [C.POINTER(ServeConfig)],
    "mr_serve_create": [_P, C.POINTER(ServeConfig), C.POINTER(_P)],
    "mr_serve_destroy": [_P],
    "mr_serve_add": [_P, _P, C.POINTER(C.c_int)],
    "mr_serve_upload_rmat": [_P, C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_int)],
    "mr_serve_upload_wait": [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _DP],
    "mr_serve_submit": [_P, C.c_int, C.c_int, C.c_uint64, _I64P],
    "mr_serve_cancel": [_P, C.c_int64],
    "mr_serve_set_plan_cache": [_P, C.c_char_p],
    "mr_serve_plan_cache_stats": [_P, _I64P, _I64P],
    "mr_serve_wait": [_P, C.c_int64, C.POINTER(C.c_int), _DP, C.POINTER(C.c_uint64)],
    "mr_serve_result": [_P, C.c_int64, _FP],
    "mr_serve_run_load": [_P, C.POINTER(ServeLoad), C.POINTER(ServeReport)],
    "mr_serve_rhs": [C.c_uint64, C.c_int64, C.c_int, _FP],
    "mr_serve_hash": [_FP, C.c_int64, C.c_int, C.POINTER(C.c_uint64)],
}
_RESTYPE = {"mr_last_error": C.c_char_p, "mr_version": C.c_char_p}

_lib = None


def lib():
    """Load libmr.so once; raise when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no fallback exists)")
        L = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != 0:
        raise MRError(code, lib().mr_last_error().decode())


def ptr(a: np.ndarray | None, ct=None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if ct is None:
        return a.ctypes.data_as(C.c_void_p)
    return a.ctypes.data_as(C.POINTER(ct))


def device_count() -> int:
    n = C.c_int(0)
    check(lib().mr_device_count(C.byref(n)))
    return n.value


# ---------------------------------------------------------------- host generators
def gen_rmat(scale: int, edge_factor: int = 16, abc=(0.57, 0.19, 0.19), seed: int = 1, scramble: bool = True,
             dedup: bool = True, symmetric: bool = False):
    """R-MAT CSR (row_ptr, col_idx) with 2^scale rows; self loops kept.
    symmetric: every edge in both directions (as symmetric MatrixMarket files
    of Graph500 R-MATs are read)."""
    m = edge_factor << scale
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    check(lib().mr_gen_rmat_edges(scale, edge_factor, abc[0], abc[1], abc[2], seed, int(scramble),
                                  ptr(src, C.c_int32), ptr(dst, C.c_int32)))
    if symmetric:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
        m = len(src)
    n = 1 << scale
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(m, np.int32)
    nnz = C.c_int64(0)
    check(lib().mr_coo_to_csr(n, m, ptr(src, C.c_int32), ptr(dst, C.c_int32), int(dedup), ptr(rp, C.c_int32),
                              ptr(ci, C.c_int32), C.byref(nnz)))
    return rp, ci[: nnz.value].copy()


def coo_to_csr(n: int, src: np.ndarray, dst: np.ndarray, dedup: bool = True):
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(max(len(src), 1), np.int32)
    nnz = C.c_int64(0)
    check(lib().mr_coo_to_csr(n, len(src), ptr(src, C.c_int32), ptr(dst, C.c_int32), int(dedup),
                              ptr(rp, C.c_int32), ptr(ci, C.c_int32), C.byref(nnz)))
    return rp, ci[: nnz.value].copy()


def gen_values(seed: int, count: int, kind: int = 0, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(max(count, 1), np.float64)
    check(lib().mr_gen_values(seed, kind, lo, hi, count, ptr(out, C.c_double)))
    return out[:count]


def gen_stencil27(nx: int, ny: int, nz: int):
    n = C.c_int64(0)
    nnz = C.c_int64(0)
    check(lib().mr_gen_stencil27(nx, ny, nz, None, None, C.byref(n), C.byref(nnz)))
    rp = np.empty(n.value + 1, np.int32)
    ci = np.empty(nnz.value, np.int32)
    check(lib().mr_gen_stencil27(nx, ny, nz, ptr(rp, C.c_int32), ptr(ci, C.c_int32), None, None))
    return rp, ci


def gen_banded(n: int, bandwidth: int, nnz_per_row: int, seed: int):
    nnz = C.c_int64(0)
    check(lib().mr_gen_banded(n, bandwidth, nnz_per_row, seed, None, None, C.byref(nnz)))
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(max(nnz.value, 1), np.int32)
    check(lib().mr_gen_banded(n, bandwidth, nnz_per_row, seed, ptr(rp, C.c_int32), ptr(ci, C.c_int32), None))
    return rp, ci[: nnz.value]


def gen_batched_set(seed: int = 3, count: int = 64):
    """The batched-launch workload: `count` matrices, rows log-uniform in
    [2^10, 2^18], nnz/row log-uniform in [4, 64]; even entries R-MAT (scale =
    round(log2 rows), edge factor = the drawn nnz/row), odd entries random
    banded (bandwidth uniform in [8, 512]).  Returns [(row_ptr, col_idx)]."""
    import math
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        rows = int(round(2.0 ** rng.uniform(10, 18)))
        d = float(2.0 ** rng.uniform(2, 6))
        if i % 2 == 0:
            scale = int(round(math.log2(rows)))
            rp, ci = gen_rmat(scale, max(1, int(round(d))), (0.57, 0.19, 0.19), seed=1000 + i)
        else:
            bw = int(rng.integers(8, 513))
            rp, ci = gen_banded(rows, bw, max(1, min(int(round(d)), 2 * bw + 1)), seed=2000 + i)
        out.append((rp, ci))
    return out


def csr_hash(rp: np.ndarray, ci: np.ndarray, n_cols: int | None = None) -> int:
    """Fingerprint of a host CSR pattern (the plan-cache key); equals
    ``Mat.hash()`` of the same matrix."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    n = len(rp) - 1
    h = C.c_uint64(0)
    check(lib().mr_csr_hash(n, n if n_cols is None else n_cols, len(ci), ptr(rp, C.c_int32), ptr(ci, C.c_int32),
                            C.byref(h)))
    return h.value


def get_width(n: int, nnz: int) -> int:
    w = C.c_int(0)
    check(lib().mr_get_width(n, nnz, C.byref(w)))
    return w.value


def check_perm(perm: np.ndarray) -> bool:
    perm = np.ascontiguousarray(perm, np.int32)
    code = lib().mr_check_perm(ptr(perm, C.c_int32), len(perm))
    if code == 4:
        return False
    check(code)
    return True


def predict(feat, model: int = MODEL_MERBIT) -> float:
    f = np.ascontiguousarray(feat, np.float64)
    out = C.c_double(0)
    check(lib().mr_predict(ptr(f, C.c_double), model, C.byref(out)))
    return out.value


def rd_params(model: int | None = None, **kw) -> RDParams:
    """Feature parameters: the reference's defaults, or those a model was
    fitted with; keyword arguments override fields."""
    p = RDParams()
    if model is None:
        check(lib().mr_rd_params_default(C.byref(p)))
    else:
        check(lib().mr_rd_params_model(model, C.byref(p)))
    for k, v in kw.items():
        if k == "bin_edges":
            p.n_edges = len(v)
            for i, e in enumerate(v):
                p.bin_edges[i] = e
        else:
            setattr(p, k, v)
    return p


def rd_exact_stream(stream: np.ndarray, n: int, params: RDParams | None = None):
    p = params or rd_params()
    s = np.ascontiguousarray(stream, np.int32)
    nb = p.n_edges + 1
    counts = np.zeros(nb, np.int64)
    feat = np.zeros(nb, np.float64)
    st = RDStats()
    check(lib().mr_rd_exact_stream(ptr(s, C.c_int32), n, len(s), C.byref(p), ptr(counts, C.c_int64),
                                   ptr(feat, C.c_double), C.byref(st)))
    return counts, feat, st.as_dict()


# ---------------------------------------------------------------- device plumbing
class DeviceArray:
    """A device buffer owned through the library's own allocator."""

    def __init__(self, ctx: "Ctx", nbytes: int, dtype=np.float32, shape=None):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = C.c_void_p()
        check(lib().mr_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p
        ctx._children.add(self)

    @classmethod
    def from_host(cls, ctx, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            check(lib().mr_memcpy(ctx.h, d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D))
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().mr_memcpy(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H))
        return out

    def zero(self):
        check(lib().mr_memset(self.ctx.h, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().mr_free(self.ctx.h, self.ptr)
            self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ctx:
    """A device context (replaces the reference's ENV, include/class.cuh:13-36)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().mr_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        self._children = weakref.WeakSet()

    def close(self):
        """Destroy the context after every server, batch, matrix and buffer
        still open on it (the library frees them against the context)."""
        if self.h:
            for kind in (Server, Batch, Mat, DeviceArray):
                for c in list(self._children):
                    if isinstance(c, kind):
                        c.close()
            lib().mr_ctx_destroy(self.h)
            self.h = None

    def arch(self):
        buf = C.create_string_buffer(64)
        cu = C.c_int(0)
        check(lib().mr_ctx_arch(self.h, buf, 64, C.byref(cu)))
        return buf.value.decode(), cu.value

    def mem_info(self):
        f = C.c_size_t(0)
        t = C.c_size_t(0)
        check(lib().mr_ctx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def stream(self):
        s = C.c_void_p()
        check(lib().mr_ctx_stream(self.h, C.byref(s)))
        return s

    def new_stream(self):
        s = C.c_void_p()
        check(lib().mr_stream_create(self.h, C.byref(s)))
        return s

    def sync(self, stream=None):
        check(lib().mr_stream_sync(self.h, stream))

    def event(self):
        e = C.c_void_p()
        check(lib().mr_event_create(self.h, C.byref(e)))
        return e

    def record(self, ev, stream=None):
        check(lib().mr_event_record(self.h, ev, stream))

    def elapsed(self, e0, e1) -> float:
        ms = C.c_float(0)
        check(lib().mr_event_elapsed_ms(self.h, e0, e1, C.byref(ms)))
        return ms.value

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int, stream=None):
        check(lib().mr_bench_copy(self.h, dst.ptr, src.ptr, nbytes, stream))

    def upload(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray.from_host(self, a)

    def empty(self, shape, dtype) -> DeviceArray:
        shape = (shape,) if isinstance(shape, int) else tuple(shape)
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return DeviceArray(self, n, dtype, shape)


class Mat:
    """A device CSR handle (mr_mat)."""

    def __init__(self, ctx: Ctx, handle):
        self.ctx = ctx
        self.h = handle
        ctx._children.add(self)

    @classmethod
    def from_csr(cls, ctx: Ctx, rp, ci, val=None, n_cols=None, store=F32):
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        n = len(rp) - 1
        nnz = int(rp[-1]) if n >= 0 else 0
        n_cols = n if n_cols is None else n_cols
        vp, vdt = None, F64
        if val is not None:
            val = np.ascontiguousarray(val)
            vdt = {np.dtype(np.float64): F64, np.dtype(np.float32): F32}[val.dtype]
            vp = val.ctypes.data_as(C.c_void_p)
        h = C.c_void_p()
        ci_arg = ci if len(ci) else np.zeros(1, np.int32)
        check(lib().mr_csr_create(ctx.h, n, n_cols, nnz, ptr(rp, C.c_int32), ptr(ci_arg, C.c_int32), vp, vdt, store,
                                  C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load_mtx(cls, ctx: Ctx, path: str, store=F32, force_ones=True):
        h = C.c_void_p()
        check(lib().mr_mtx_load(ctx.h, str(path).encode(), store, int(force_ones), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load_gap(cls, ctx: Ctx, path: str, symmetrize: bool = False, store=F32):
        """A GAP-benchmark graph file (.el .net .wel .gr .graph .sg .wsg)."""
        h = C.c_void_p()
        check(lib().mr_gap_load(ctx.h, str(path).encode(), int(symmetrize), store, C.byref(h)))
        return cls(ctx, h)

    def close(self):
        if self.h:
            lib().mr_mat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        r, c, z, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        check(lib().mr_mat_info(self.h, C.byref(r), C.byref(c), C.byref(z), C.byref(d)))
        return r.value, c.value, z.value, d.value

    @property
    def vec_dtype(self):
        return np.float64 if self.info()[3] == F64 else np.float32

    def export(self):
        n, _, nnz, dt = self.info()
        rp = np.empty(n + 1, np.int32)
        ci = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1), _NP_OF[dt])
        check(lib().mr_mat_export(self.h, ptr(rp, C.c_int32), ptr(ci, C.c_int32), ptr(val)))
        return rp, ci[:nnz], val[:nnz]

    def convert(self, dtype: int) -> "Mat":
        h = C.c_void_p()
        check(lib().mr_mat_convert(self.h, dtype, C.byref(h)))
        return Mat(self.ctx, h)

    def perm(self) -> np.ndarray:
        n = self.info()[0]
        p = np.empty(max(n, 1), np.int32)
        check(lib().mr_mat_perm(self.h, ptr(p, C.c_int32)))
        return p[:n]

    def permute(self, perm):
        perm = np.ascontiguousarray(perm, np.int32)
        check(lib().mr_permute(self.h, ptr(perm, C.c_int32)))

    def permute_vec(self, direction: int, src: DeviceArray, dst: DeviceArray, stream=None):
        check(lib().mr_permute_vec(self.h, direction, src.ptr, dst.ptr, stream))

    # reorder
    def reorder_pgorder(self, width: int = 0, window: int = 5):
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        check(lib().mr_reorder_pgorder(self.h, width, window, ptr(perm, C.c_int32), C.byref(st)))
        return perm[:n], st.as_dict()

    def pgorder_order(self, width: int = 0, window: int = 5):
        n = self.info()[0]
        g = np.empty(max(n, 1), np.uint32)
        f = np.empty(max(n, 1), np.uint32)
        st = ReorderStats()
        check(lib().mr_pgorder_order(self.h, width, window, ptr(g, C.c_uint32), ptr(f, C.c_uint32), C.byref(st)))
        return g[:n], f[:n], st.as_dict()

    def pgorder_cpu_phase(self, gpu_order: np.ndarray, width: int = 0, window: int = 5, threads: int = 0):
        """gorder6's CPU phase alone on `gpu_order` (mr_pgorder_order's GPU
        placement): (final_order, wall ms)."""
        n = self.info()[0]
        g = np.ascontiguousarray(gpu_order, np.uint32)
        if len(g) != n:
            raise ValueError("gpu_order must hold one entry per row")
        f = np.empty(max(n, 1), np.uint32)
        ms = C.c_double()
        check(lib().mr_pgorder_cpu_phase(self.h, width, window, ptr(g, C.c_uint32), ptr(f, C.c_uint32), threads,
                                         C.byref(ms)))
        return f[:n], ms.value

    def reorder_gorder5(self, vsize: int = 0, width1: int = 1, width2: int = 5):
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        check(lib().mr_reorder_gorder5(self.h, vsize, width1, width2, ptr(perm, C.c_int32), C.byref(st)))
        return perm[:n], st.as_dict()

    def gorder5_order(self, vsize: int = 0, width1: int = 1, width2: int = 5):
        n = self.info()[0]
        g = np.empty(max(n, 1), np.uint32)
        f = np.empty(max(n, 1), np.uint32)
        st = ReorderStats()
        check(lib().mr_gorder5_order(self.h, vsize, width1, width2, ptr(g, C.c_uint32), ptr(f, C.c_uint32),
                                     C.byref(st)))
        return g[:n], f[:n], st.as_dict()

    def reorder_tile_order(self, width1: int = 1024, width2: int = 512):
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        check(lib().mr_reorder_tile_order(self.h, width1, width2, ptr(perm, C.c_int32), C.byref(st)))
        return perm[:n], st.as_dict()

    def reorder_degree(self):
        return self._reorder_plain(lib().mr_reorder_degree)

    def reorder_dbg(self):
        return self._reorder_plain(lib().mr_reorder_dbg)

    def _reorder_plain(self, fn):
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        check(fn(self.h, ptr(perm, C.c_int32), C.byref(st)))
        return perm[:n], st.as_dict()

    def reorder_pgorder_cached(self, cache_dir: str, width: int = 0, window: int = 5):
        """pGOrder through the on-disk plan cache; returns (perm, stats, hit)
        with hit 1 = entry applied, 0 = computed and stored, 2 = a stale entry
        replaced, -1 = computed but not stored."""
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        hit = C.c_int(0)
        check(lib().mr_reorder_pgorder_cached(self.h, width, window, str(cache_dir).encode(),
                                              ptr(perm, C.c_int32), C.byref(st), C.byref(hit)))
        return perm[:n], st.as_dict(), hit.value

    def plan_save(self):
        check(lib().mr_plan_save(self.h))

    def hash(self) -> int:
        h = C.c_uint64(0)
        check(lib().mr_mat_hash(self.h, C.byref(h)))
        return h.value

    def reorder_gorder_cpu(self, window: int = 5):
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        check(lib().mr_reorder_gorder_cpu(self.h, window, ptr(perm, C.c_int32), C.byref(st)))
        return perm[:n], st.as_dict()

    def reorder_rcm(self):
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        check(lib().mr_reorder_rcm(self.h, ptr(perm, C.c_int32), C.byref(st)))
        return perm[:n], st.as_dict()

    def reorder_rabbit(self, reference_scatter: bool = False):
        n = self.info()[0]
        perm = np.empty(max(n, 1), np.int32)
        st = ReorderStats()
        check(lib().mr_reorder_rabbit(self.h, int(reference_scatter), ptr(perm, C.c_int32), C.byref(st)))
        return perm[:n], st.as_dict()

    # reuse distance
    def rd_features(self, mode=RD_SAMPLED, params: RDParams | None = None):
        p = params or rd_params()
        feat = np.zeros(p.n_edges + 1, np.float64)
        st = RDStats()
        check(lib().mr_rd_features(self.h, mode, C.byref(p), ptr(feat, C.c_double), C.byref(st)))
        return feat, st.as_dict()

    def rd_sampled_raw(self, params: RDParams | None = None):
        p = params or rd_params()
        s = p.samples
        est = np.zeros(s, np.float32)
        status = np.zeros(s, np.int32)
        pos = np.zeros(s, np.int32)
        check(lib().mr_rd_sampled_raw(self.h, C.byref(p), ptr(est, C.c_float), ptr(status, C.c_int32),
                                      ptr(pos, C.c_int32)))
        return est, status, pos

    def locality_predict(self, model=MODEL_MERBIT):
        sp = C.c_double(0)
        feat = np.zeros(10, np.float64)
        st = RDStats()
        check(lib().mr_locality_predict(self.h, model, C.byref(sp), ptr(feat, C.c_double), C.byref(st)))
        return sp.value, feat, st.as_dict()

    # multiply
    def spmv(self, x: DeviceArray, y: DeviceArray, kernel=MERBIT, stream=None):
        check(lib().mr_spmv(self.h, kernel, x.ptr, y.ptr, stream))

    def spmv_ws_bytes(self) -> int:
        b = C.c_size_t(0)
        check(lib().mr_merbit_ws_bytes(self.h, C.byref(b)))
        return b.value

    def spmv_ws(self, x: DeviceArray, y: DeviceArray, ws: DeviceArray, stream=None):
        check(lib().mr_spmv_ws(self.h, x.ptr, y.ptr, ws.ptr, stream))

    def spmv_noalloc(self, x: DeviceArray, y: DeviceArray, stream=None):
        """MERBIT SpMV with the carries in the library's static scratch array."""
        check(lib().mr_spmv_noalloc(self.h, x.ptr, y.ptr, stream))

    def spmm(self, k: int, X: DeviceArray, Y: DeviceArray, kernel=MERBIT, ldx=None, ldy=None, stream=None):
        check(lib().mr_spmm(self.h, kernel, k, X.ptr, ldx or k, Y.ptr, ldy or k, stream))

    def merbit_configure(self, sigma: int):
        check(lib().mr_merbit_configure(self.h, sigma))

    def merbit_options(self, policy: int):
        check(lib().mr_merbit_options(self.h, policy))

    def merbit_policy(self) -> int:
        p = C.c_int(0)
        check(lib().mr_merbit_get_options(self.h, C.byref(p)))
        return p.value

    def merbit_autotune(self, reps: int = 5):
        sg, pol, us = C.c_int(), C.c_int(), C.c_float()
        check(lib().mr_merbit_autotune(self.h, reps, C.byref(sg), C.byref(pol), C.byref(us)))
        return {"sigma": sg.value, "policy": pol.value, "us": us.value}

    def merbit_xc_configure(self, items_per_lane: int = 16, waves: int = 16, cached_cols: int = -1):
        check(lib().mr_merbit_xc_configure(self.h, items_per_lane, waves, cached_cols))

    def merbit_xc_info(self):
        s, w, c, t = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        check(lib().mr_merbit_xc_info(self.h, C.byref(s), C.byref(w), C.byref(c), C.byref(t)))
        return {"items_per_lane": s.value, "waves": w.value, "cached_cols": c.value, "tiles": t.value}

    def merbit_panel_configure(self, items_per_lane: int = 8, waves: int = 16):
        check(lib().mr_merbit_panel_configure(self.h, items_per_lane, waves))

    def merbit_panel_info(self):
        s, w, wg = C.c_int(), C.c_int(), C.c_int()
        t, pa, sg, ms = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        check(lib().mr_merbit_panel_info(self.h, C.byref(s), C.byref(w), C.byref(t), C.byref(pa), C.byref(sg),
                                         C.byref(wg), C.byref(ms)))
        return {"items_per_lane": s.value, "waves": w.value, "tiles": t.value, "partials": pa.value,
                "segments": sg.value, "workgroups": wg.value, "build_ms": ms.value}

    def merbit_stored_bytes(self) -> int:
        """Bytes one SpMV moves in the current variant's stored format (-1
        before its plan is built)."""
        b = C.c_int64(-1)
        check(lib().mr_merbit_stored_bytes(self.h, C.byref(b)))
        return b.value

    def merbit_info(self):
        s, it, nb, ms = C.c_int(), C.c_int64(), C.c_int64(), C.c_double()
        check(lib().mr_merbit_info(self.h, C.byref(s), C.byref(it), C.byref(nb), C.byref(ms)))
        return {"sigma": s.value, "items": it.value, "n_blocks": nb.value, "build_ms": ms.value}

    def merbit_export(self):
        info = self.merbit_info()
        words = info["n_blocks"] * 256 * info["sigma"] // 32
        flags = np.zeros(max(words, 1), np.uint32)
        bx = np.zeros(info["n_blocks"] + 1, np.int32)
        by = np.zeros(info["n_blocks"] + 1, np.int32)
        check(lib().mr_merbit_export(self.h, ptr(flags, C.c_uint32), ptr(bx, C.c_int32), ptr(by, C.c_int32)))
        return flags[:words], bx, by, info


class Batch:
    """Many matrices multiplied in one persistent fused launch (mr_batch_*)."""

    def __init__(self, ctx: Ctx, mats, k: int = 1):
        self.ctx = ctx
        self.mats = list(mats)
        self.k = k
        arr = (_P * len(self.mats))(*[m.h for m in self.mats])
        h = _P()
        check(lib().mr_batch_create(ctx.h, len(self.mats), arr, k, C.byref(h)))
        self.h = h
        ctx._children.add(self)

    def close(self):
        if getattr(self, "h", None):
            lib().mr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_io(self, xs, ys, ldx=None, ldy=None, stream=None):
        n = len(self.mats)
        xa = (_P * n)(*[x.ptr if x is not None else None for x in xs])
        ya = (_P * n)(*[y.ptr if y is not None else None for y in ys])
        lx = np.asarray(ldx, np.int64) if ldx is not None else None
        ly = np.asarray(ldy, np.int64) if ldy is not None else None
        check(lib().mr_batch_set_io(self.h, xa, ptr(lx, C.c_int64), ya, ptr(ly, C.c_int64), stream))
        self._io = (list(xs), list(ys))   # keep the buffers alive

    def configure(self, sigma: int = 8, nt: int = 1, grid: int = 0):
        check(lib().mr_batch_configure(self.h, sigma, nt, grid))

    def autotune(self, reps: int = 5):
        s, t, g, us = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        check(lib().mr_batch_autotune(self.h, reps, C.byref(s), C.byref(t), C.byref(g), C.byref(us)))
        return {"sigma": s.value, "nt": t.value, "grid": g.value, "us": us.value}

    def run(self, stream=None):
        check(lib().mr_batch_run(self.h, stream))

    def capture(self, stream):
        check(lib().mr_batch_graph_capture(self.h, stream))

    def launch(self, stream=None):
        check(lib().mr_batch_graph_launch(self.h, stream))

    def info(self):
        t, z, b, f = C.c_int64(), C.c_int64(), C.c_double(), C.c_double()
        check(lib().mr_batch_info(self.h, C.byref(t), C.byref(z), C.byref(b), C.byref(f)))
        return {"tiles": t.value, "nnz": z.value, "bytes": b.value, "flops": f.value}


def serve_rhs(seed: int, n_cols: int, k: int) -> np.ndarray:
    """Host copy of the serving loop's right-hand sides: X[n_cols][k]."""
    out = np.empty((max(n_cols, 0), k), np.float32)
    check(lib().mr_serve_rhs(seed, n_cols, k, ptr(out, C.c_float)))
    return out


def serve_hash(y: np.ndarray) -> int:
    """Host copy of the serving loop's result hash of Y[n_rows][k]."""
    y = np.ascontiguousarray(y, np.float32)
    y2 = y.reshape(y.shape[0], -1) if y.ndim > 1 else y.reshape(-1, 1)
    h = C.c_uint64(0)
    check(lib().mr_serve_hash(ptr(y2, C.c_float), y2.shape[0], y2.shape[1], C.byref(h)))
    return h.value


class Server:
    """The serving loop (mr_serve_*)."""

    def __init__(self, ctx: Ctx, **cfg):
        self.ctx = ctx
        c = ServeConfig()
        check(lib().mr_serve_config_default(C.byref(c)))
        for key, v in cfg.items():
            setattr(c, key, v)
        self.cfg = c
        h = _P()
        check(lib().mr_serve_create(ctx.h, C.byref(c), C.byref(h)))
        self.h = h
        self._mats = []
        ctx._children.add(self)

    def close(self):
        if getattr(self, "h", None):
            lib().mr_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, mat: "Mat") -> int:
        i = C.c_int(-1)
        check(lib().mr_serve_add(self.h, mat.h, C.byref(i)))
        self._mats.append(mat)
        return i.value

    def upload_rmat(self, scale: int, edge_factor: int = 16, seed: int = 1) -> int:
        t = C.c_int(-1)
        check(lib().mr_serve_upload_rmat(self.h, scale, edge_factor, seed, C.byref(t)))
        return t.value

    def upload_wait(self, ticket: int):
        st, mid, ms = C.c_int(), C.c_int(), C.c_double()
        check(lib().mr_serve_upload_wait(self.h, ticket, C.byref(st), C.byref(mid), C.byref(ms)))
        return st.value, mid.value, ms.value

    def submit(self, mat_id: int, k: int, seed: int) -> int:
        r = C.c_int64(-1)
        check(lib().mr_serve_submit(self.h, mat_id, k, seed, C.byref(r)))
        return r.value

    def cancel(self, req: int):
        check(lib().mr_serve_cancel(self.h, req))

    def set_plan_cache(self, cache_dir: str | None):
        check(lib().mr_serve_set_plan_cache(self.h, (str(cache_dir) if cache_dir else "").encode()))

    def plan_cache_stats(self):
        h, m = C.c_int64(0), C.c_int64(0)
        check(lib().mr_serve_plan_cache_stats(self.h, C.byref(h), C.byref(m)))
        return h.value, m.value

    def wait(self, req: int):
        st, lat, h = C.c_int(), C.c_double(), C.c_uint64()
        check(lib().mr_serve_wait(self.h, req, C.byref(st), C.byref(lat), C.byref(h)))
        return st.value, lat.value, h.value

    def result(self, req: int, n_rows: int, k: int) -> np.ndarray:
        y = np.empty((n_rows, k), np.float32)
        check(lib().mr_serve_result(self.h, req, ptr(y, C.c_float)))
        return y

    def run_load(self, rate: float, n_requests: int, seed: int = 1, ks=(1, 4, 16), upload_every: int = 0,
                 upload_scale: int = 18, upload_edge_factor: int = 16) -> dict:
        ld = ServeLoad()
        ld.rate, ld.n_requests, ld.seed = rate, n_requests, seed
        for i, k in enumerate(ks):
            ld.k_choices[i] = k
        ld.n_k = len(ks)
        ld.upload_every, ld.upload_scale, ld.upload_edge_factor = upload_every, upload_scale, upload_edge_factor
        rep = ServeReport()
        check(lib().mr_serve_run_load(self.h, C.byref(ld), C.byref(rep)))
        return rep.as_dict()


class Graph:
    """C++-style mirror of the reference's GRAPH<V> workflow (include/class.cuh:59-99):
    load, locality_predict, gorder6 / gorder1 / relabel_rcm, spmv."""

    def __init__(self, ctx: Ctx, mat: Mat):
        self.ctx = ctx
        self.mat = mat

    def get_width(self) -> int:
        n, _, nnz, _ = self.mat.info()
        return get_width(n, nnz)

    def gorder6(self, width: int = 0):
        return self.mat.reorder_pgorder(width or self.get_width(), 5)

    def gorder1(self, width: int = 5):
        return self.mat.reorder_gorder_cpu(width)

    def relabel_rcm(self):
        return self.mat.reorder_rcm()

    def relabel_degree(self):
        return self.mat.reorder_degree()

    def gorder5(self, vsize: int = 0, width1: int = 1, width2: int = 5):
        return self.mat.reorder_gorder5(vsize, width1, width2)

    def order(self, width1: int, width2: int):
        return self.mat.reorder_tile_order(width1, width2)

    def dbg(self):
        return self.mat.reorder_dbg()

    def locality_predict(self) -> float:
        return self.mat.locality_predict(MODEL_MERBIT)[0]

    def spmv(self, x: np.ndarray, method: str = "merbit") -> np.ndarray:
        """y = A x on the current (relabelled) matrix, x and y in the ORIGINAL
        vertex order (x is relabelled on the way in, y on the way out, as
        class.cu:1437-1445 does for its check)."""
        kernel = {"merbit": MERBIT, "csr": CSR_VECTOR, "csr_scalar": CSR_SCALAR}[method]
        dt = self.mat.vec_dtype
        xd = self.ctx.upload(np.ascontiguousarray(x, dt))
        xp = self.ctx.empty(len(x), dt)
        self.mat.permute_vec(0, xd, xp)
        yp = self.ctx.empty(self.mat.info()[0], dt)
        self.mat.spmv(xp, yp, kernel)
        y = self.ctx.empty(self.mat.info()[0], dt)
        self.mat.permute_vec(1, yp, y)
        return y.to_host()
