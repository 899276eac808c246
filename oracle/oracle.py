"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's SpMV locality pipeline
(see oracle.h).  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py may import this module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_DIR = HERE / "_ref"

_I32P = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_U8P = C.POINTER(C.c_uint8)

_SIG = {
    "oracle_heap_trace": (C.c_int64, [C.c_int32, C.c_int64, _I32P, C.c_int64, _I32P]),
    "oracle_get_width": (C.c_int32, [C.c_int64, C.c_int64]),
    "oracle_transpose": (None, [C.c_int32, C.c_int32, _I32P, _I32P, _I32P, _I32P]),
    "oracle_check_perm": (C.c_int, [_I32P, C.c_int64]),
    "oracle_permute": (None, [C.c_int32, _I32P, _I32P, _DP, _I32P, _I32P, _I32P, _DP]),
    "oracle_order_degree": (None, [C.c_int32, _I32P, _I32P, _I32P]),
    "oracle_tile_order": (C.c_int, [C.c_int32, _I32P, _I32P, C.c_int32, C.c_int32, _U32P, _I32P]),
    "oracle_gorder5": (C.c_int, [C.c_int32, _I32P, _I32P, C.c_int32, C.c_int32, C.c_int32, _U32P, _U32P, _I32P]),
    "oracle_order_dbg": (None, [C.c_int32, _I32P, _I32P, _I32P]),
    "oracle_pgorder_schedule": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, _I32P, _I32P, _I32P, _I32P]),
    "oracle_pgorder_gpu_loop": (C.c_int, [C.c_int32, _I32P, _I32P, _I32P, _I32P, C.c_int32, _U32P]),
    "oracle_hll_hash_bias": (C.c_double, [C.c_int32, C.c_uint64]),
    "oracle_pgorder_gpu_loop_ex": (C.c_int, [C.c_int32, _I32P, _I32P, _I32P, _I32P, C.c_int32, C.c_int32, C.c_int32,
                                             _U32P]),
    "oracle_order_on_cpu": (None, [C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P, _U32P, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_int32]),
    "oracle_pgorder": (C.c_int, [C.c_int32, _I32P, _I32P, C.c_int32, C.c_int32, _U32P, _U32P, _I32P]),
    "oracle_gorder1": (C.c_int, [C.c_int32, _I32P, _I32P, C.c_int32, _U32P, _I32P]),
    "oracle_gorder1_score": (C.c_int64, [C.c_int32, _I32P, _I32P, _I32P, _I32P, _U32P, C.c_int32, C.c_int32,
                                         C.c_int32]),
    "oracle_rcm": (C.c_int, [C.c_int32, C.c_int64, _I32P, _I32P, _U32P]),
    "oracle_rabbit": (C.c_int, [C.c_int32, _I32P, _I32P, _U32P]),
    "oracle_rd_exact": (C.c_int, [_I32P, C.c_int32, C.c_int64, C.c_int32, _I32P, C.c_int32, _I64P, _I64P, _DP]),
    "oracle_rd_bruteforce": (C.c_int, [_I32P, C.c_int64, C.c_int32, _I64P]),
    "oracle_hll_hash": (C.c_uint32, [C.c_uint32]),
    "oracle_hll_rank": (C.c_int32, [C.c_uint32, C.c_int32]),
    "oracle_hll_estimate": (C.c_float, [_I32P, C.c_int32]),
    "oracle_hll_sample": (None, [_I32P, C.c_int64, C.c_int32, C.c_int32, C.c_uint32, C.c_int32, C.c_int32,
                                 C.c_int32, _FP, _I32P, _I32P]),
    "oracle_hll_seed": (C.c_uint32, [C.c_int32, C.c_int32]),
    "oracle_rd_sampled": (C.c_int, [_I32P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_uint32, _I32P, C.c_int32, _FP, _I32P, _I32P, _I64P, _I32P, _DP]),
    "oracle_predict": (C.c_double, [_DP, C.c_int32]),
    "oracle_predict_preprocessed": (C.c_double, [_DP, _U8P, C.c_int32]),
    "oracle_spmv_f64": (None, [C.c_int32, _I32P, _I32P, _DP, _DP, _DP, _DP]),
    "oracle_spmm_f64": (None, [C.c_int32, C.c_int32, _I32P, _I32P, _DP, _DP, _DP]),
    "oracle_spmv_f32_threads": (C.c_double, [C.c_int32, _I32P, _I32P, _FP, _FP, _FP, C.c_int32, C.c_int32]),
    "oracle_spmv_f64_threads": (C.c_double, [C.c_int32, _I32P, _I32P, _DP, _DP, _DP, C.c_int32, C.c_int32]),
    "oracle_merbit_ref_tiles": (C.c_int, [C.c_int32, C.c_int64, _I32P, C.c_int32, _U32P, _U32P, _U32P]),
    "oracle_merge_search": (None, [_I32P, C.c_int32, C.c_int64, C.c_int64, _I64P, _I64P]),
    "oracle_merbit_spmv": (None, [C.c_int32, C.c_int64, _I32P, _I32P, _DP, _DP, C.c_int32, _DP]),
}

_lib = None

DEFAULT_EDGES = np.array([1, 112, 512, 18432, 36864, 73728, 147456, 294912, 589824], np.int32)


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def transpose(n_rows, n_cols, rp, ci):
    rp, ci = _i32(rp), _i32(ci)
    tp = np.empty(n_cols + 1, np.int32)
    ti = np.empty(max(len(ci), 1), np.int32)
    lib().oracle_transpose(n_rows, n_cols, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(tp, C.c_int32),
                           _p(ti, C.c_int32))
    return tp, ti[: len(ci)]


def heap_trace(n_nodes, buckets, ops):
    ops = _i32(np.asarray(ops).reshape(-1))
    out = np.empty(len(ops) // 3 + 1, np.int32)
    k = lib().oracle_heap_trace(n_nodes, buckets, _p(ops, C.c_int32), len(ops) // 3, _p(out, C.c_int32))
    return out[:k]


def get_width(n, nnz):
    return lib().oracle_get_width(n, nnz)


def check_perm(perm):
    perm = _i32(perm)
    return lib().oracle_check_perm(_p(perm, C.c_int32), len(perm)) == 0


def gorder5(n, rp, ci, vsize, width1=1, width2=5):
    """(gpu placement, final placement, perm) of GRAPH::gorder5."""
    rp, ci = _i32(rp), _i32(ci if len(ci) else np.zeros(1, np.int32))
    g = np.zeros(max(n, 1), np.uint32)
    f = np.zeros(max(n, 1), np.uint32)
    perm = np.zeros(max(n, 1), np.int32)
    lib().oracle_gorder5(n, _p(rp, C.c_int32), _p(ci, C.c_int32), vsize, width1, width2, _p(g, C.c_uint32),
                         _p(f, C.c_uint32), _p(perm, C.c_int32))
    return g[:n], f[:n], perm[:n]


def tile_order(n, rp, ci, width1, width2):
    """(order, perm) of GRAPH::order."""
    rp, ci = _i32(rp), _i32(ci if len(ci) else np.zeros(1, np.int32))
    order = np.zeros(max(n, 1), np.uint32)
    perm = np.zeros(max(n, 1), np.int32)
    lib().oracle_tile_order(n, _p(rp, C.c_int32), _p(ci, C.c_int32), width1, width2, _p(order, C.c_uint32),
                            _p(perm, C.c_int32))
    return order[:n], perm[:n]


def order_degree(n, rp, ci, grouped=False):
    """perm[old] = new of the degree order (grouped=False) or DBG (True)."""
    rp, ci = _i32(rp), _i32(ci if len(ci) else np.zeros(1, np.int32))
    perm = np.zeros(max(n, 1), np.int32)
    fn = lib().oracle_order_dbg if grouped else lib().oracle_order_degree
    fn(n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(perm, C.c_int32))
    return perm[:n]


def permute(n, rp, ci, val, perm):
    rp, ci, perm = _i32(rp), _i32(ci), _i32(perm)
    val = np.ascontiguousarray(val if val is not None else np.ones(len(ci)), np.float64)
    rp2 = np.empty(n + 1, np.int32)
    ci2 = np.empty(max(len(ci), 1), np.int32)
    v2 = np.empty(max(len(ci), 1), np.float64)
    lib().oracle_permute(n, _p(rp, C.c_int32), _p(ci if len(ci) else np.zeros(1, np.int32), C.c_int32),
                         _p(val if len(val) else np.zeros(1), C.c_double), _p(perm, C.c_int32), _p(rp2, C.c_int32),
                         _p(ci2, C.c_int32), _p(v2, C.c_double))
    return rp2, ci2[: len(ci)], v2[: len(ci)]


def pgorder_schedule(n, width):
    cap = (n + width - 1) // width + 64
    a = [np.zeros(cap, np.int32) for _ in range(4)]
    b = lib().oracle_pgorder_schedule(n, width, cap, *[_p(x, C.c_int32) for x in a])
    return [x[:b] for x in a]


def pgorder_gpu_loop(n, rp, ci, width, hub_cap=10000, max_batches=-1, tp=None, ti=None):
    """Placement (position -> vertex) after the GPU loop; with max_batches >= 0
    only the positions of the first max_batches batches are filled (the rest
    stay 0xFFFFFFFF)."""
    rp, ci = _i32(rp), _i32(ci)
    if tp is None:
        tp, ti = transpose(n, n, rp, ci)
    pl = np.full(n, 0xFFFFFFFF, np.uint32)
    lib().oracle_pgorder_gpu_loop_ex(n, _p(rp, C.c_int32), _p(ci if len(ci) else np.zeros(1, np.int32), C.c_int32),
                                     _p(_i32(tp), C.c_int32),
                                     _p(_i32(ti) if len(ti) else np.zeros(1, np.int32), C.c_int32), width, hub_cap,
                                     max_batches, _p(pl, C.c_uint32))
    return pl


def pgorder(n, rp, ci, width=1024, window=5):
    rp, ci = _i32(rp), _i32(ci)
    g = np.empty(n, np.uint32)
    f = np.empty(n, np.uint32)
    perm = np.empty(n, np.int32)
    lib().oracle_pgorder(n, _p(rp, C.c_int32), _p(ci if len(ci) else np.zeros(1, np.int32), C.c_int32), width,
                         window, _p(g, C.c_uint32), _p(f, C.c_uint32), _p(perm, C.c_int32))
    return g, f, perm


def order_on_cpu(n, rp, ci, slice_, batch_id, window, degree, tp=None, ti=None, batch_of=None):
    rp, ci = _i32(rp), _i32(ci)
    if tp is None:
        tp, ti = transpose(n, n, rp, ci)
    tp, ti = _i32(tp), _i32(ti)
    if batch_of is None:
        batch_of = np.full(n, -1, np.int32)
    s = np.ascontiguousarray(slice_, np.uint32).copy()
    lib().oracle_order_on_cpu(n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(tp, C.c_int32), _p(ti, C.c_int32),
                              _p(batch_of, C.c_int32), _p(s, C.c_uint32), batch_id, len(s), window, degree)
    return s


def gorder1(n, rp, ci, window=5):
    rp, ci = _i32(rp), _i32(ci)
    order = np.empty(n, np.uint32)
    perm = np.empty(n, np.int32)
    lib().oracle_gorder1(n, _p(rp, C.c_int32), _p(ci if len(ci) else np.zeros(1, np.int32), C.c_int32), window,
                         _p(order, C.c_uint32), _p(perm, C.c_int32))
    return order, perm


def gorder1_score(n, rp, ci, tp, ti, order, pid, window, z):
    return lib().oracle_gorder1_score(n, _p(_i32(rp), C.c_int32), _p(_i32(ci), C.c_int32), _p(_i32(tp), C.c_int32),
                                      _p(_i32(ti), C.c_int32), _p(np.ascontiguousarray(order, np.uint32), C.c_uint32),
                                      pid, window, z)


def rcm(n, rows, cols):
    rows, cols = _i32(rows), _i32(cols)
    out = np.empty(n, np.uint32)
    lib().oracle_rcm(n, len(rows), _p(rows, C.c_int32), _p(cols, C.c_int32), _p(out, C.c_uint32))
    return out


def rabbit(n, rp, ci):
    """Rabbit Order in one deterministic order: (new_id per vertex, communities)."""
    rp, ci = _i32(rp), _i32(ci)
    out = np.empty(max(n, 1), np.uint32)
    ncom = lib().oracle_rabbit(n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(out, C.c_uint32))
    return out[:n], ncom


def rd_exact(stream, n, sector=32, edges=DEFAULT_EDGES):
    s, e = _i32(stream), _i32(edges)
    counts = np.zeros(len(e) + 1, np.int64)
    prob = np.zeros(len(e) + 1, np.float64)
    reuses = C.c_int64(0)
    lib().oracle_rd_exact(_p(s, C.c_int32), n, len(s), sector, _p(e, C.c_int32), len(e), _p(counts, C.c_int64),
                          C.byref(reuses), _p(prob, C.c_double))
    return counts, reuses.value, prob


def rd_bruteforce(stream, sector):
    s = _i32(stream)
    out = np.empty(len(s), np.int64)
    lib().oracle_rd_bruteforce(_p(s, C.c_int32), len(s), sector, _p(out, C.c_int64))
    return out


def hll_hash(x):
    return lib().oracle_hll_hash(x)


def hll_hash_bias(log2_n=23, seed=1):
    return lib().oracle_hll_hash_bias(log2_n, seed)


def hll_rank(h, p):
    return lib().oracle_hll_rank(h, p)


def hll_estimate(regs):
    r = _i32(regs)
    return lib().oracle_hll_estimate(_p(r, C.c_int32), len(r))


def hll_seed(graph_id=0, stream_type=0):
    return lib().oracle_hll_seed(graph_id, stream_type)


def rd_sampled(stream, samples=30960, sector=32, rd_cap=589824, hll_p=12, max_scan=0, seed=None,
               edges=DEFAULT_EDGES):
    s, e = _i32(stream), _i32(edges)
    seed = hll_seed() if seed is None else seed
    est = np.zeros(samples, np.float32)
    st = np.zeros(samples, np.int32)
    pos = np.zeros(samples, np.int32)
    counts = np.zeros(len(e) + 1, np.int64)
    stat4 = np.zeros(4, np.int32)
    prob = np.zeros(len(e) + 1, np.float64)
    lib().oracle_rd_sampled(_p(s, C.c_int32), len(s), samples, sector, rd_cap, hll_p, max_scan, seed,
                            _p(e, C.c_int32), len(e), _p(est, C.c_float), _p(st, C.c_int32), _p(pos, C.c_int32),
                            _p(counts, C.c_int64), _p(stat4, C.c_int32), _p(prob, C.c_double))
    return {"est": est, "status": st, "pos": pos, "counts": counts, "stat4": stat4, "prob": prob}


def predict(feat, model=0):
    f = np.ascontiguousarray(feat, np.float64)
    return lib().oracle_predict(_p(f, C.c_double), model)


def predict_preprocessed(feat, missing):
    f = np.ascontiguousarray(feat, np.float64)
    m = np.ascontiguousarray(missing, np.uint8)
    return lib().oracle_predict_preprocessed(_p(f, C.c_double), _p(m, C.c_uint8), 0)


def spmv_f64(rp, ci, val, x):
    rp, ci = _i32(rp), _i32(ci)
    n = len(rp) - 1
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(max(n, 1), np.float64)
    a = np.empty(max(n, 1), np.float64)
    ci_ = ci if len(ci) else np.zeros(1, np.int32)
    val_ = val if len(val) else np.zeros(1)
    x_ = x if len(x) else np.zeros(1)
    lib().oracle_spmv_f64(n, _p(rp, C.c_int32), _p(ci_, C.c_int32), _p(val_, C.c_double), _p(x_, C.c_double),
                          _p(y, C.c_double), _p(a, C.c_double))
    return y[:n], a[:n]


def spmm_f64(rp, ci, val, X):
    rp, ci = _i32(rp), _i32(ci)
    n = len(rp) - 1
    X = np.ascontiguousarray(X, np.float64)
    k = X.shape[1]
    Y = np.empty((n, k), np.float64)
    lib().oracle_spmm_f64(n, k, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(np.ascontiguousarray(val, np.float64),
                          C.c_double), _p(X, C.c_double), _p(Y, C.c_double))
    return Y


def spmv_f32_threads(rp, ci, val, x, threads, iters):
    rp, ci = _i32(rp), _i32(ci)
    n = len(rp) - 1
    val = np.ascontiguousarray(val, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(n, np.float32)
    ms = lib().oracle_spmv_f32_threads(n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(val, C.c_float),
                                       _p(x, C.c_float), _p(y, C.c_float), threads, iters)
    return y, ms


def spmv_f64_threads(rp, ci, val, x, threads, iters):
    """FP64 twin of spmv_f32_threads (the CPU baseline's second precision)."""
    rp, ci = _i32(rp), _i32(ci)
    n = len(rp) - 1
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(n, np.float64)
    ms = lib().oracle_spmv_f64_threads(n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(val, C.c_double),
                                       _p(x, C.c_double), _p(y, C.c_double), threads, iters)
    return y, ms


def merbit_ref_tiles(rp, nnz, sigma=14):
    rp = _i32(rp)
    n = len(rp) - 1
    lanes = (n + nnz + sigma - 1) // sigma
    tiles = (lanes + 31) // 32
    tx = np.zeros(tiles + 1, np.uint32)
    ty = np.zeros(tiles + 1, np.uint32)
    desc = np.zeros(max(lanes, 1), np.uint32)
    lib().oracle_merbit_ref_tiles(n, nnz, _p(rp, C.c_int32), sigma, _p(tx, C.c_uint32), _p(ty, C.c_uint32),
                                  _p(desc, C.c_uint32))
    return tx, ty, desc[:lanes]


def merge_search(rp, nnz, diag):
    rp = _i32(rp)
    cx, cy = C.c_int64(0), C.c_int64(0)
    lib().oracle_merge_search(_p(rp, C.c_int32), len(rp) - 1, nnz, diag, C.byref(cx), C.byref(cy))
    return cx.value, cy.value


def merbit_spmv(rp, ci, val, x, sigma=14):
    rp, ci = _i32(rp), _i32(ci)
    n = len(rp) - 1
    y = np.empty(n, np.float64)
    lib().oracle_merbit_spmv(n, int(rp[-1]), _p(rp, C.c_int32), _p(ci, C.c_int32),
                             _p(np.ascontiguousarray(val, np.float64), C.c_double),
                             _p(np.ascontiguousarray(x, np.float64), C.c_double), sigma, _p(y, C.c_double))
    return y


def spmv_parity(y, yhat, absrow):
    """Condition-aware parity metrics: max_i |y-yhat|/absrow_i and relative L2."""
    y = np.asarray(y, np.float64)
    yhat = np.asarray(yhat, np.float64)
    absrow = np.asarray(absrow, np.float64)
    d = np.abs(y - yhat)
    denom = np.where(absrow > 0, absrow, 1.0)
    rel = np.where(absrow > 0, d / denom, d)
    n2 = np.linalg.norm(yhat)
    return float(rel.max(initial=0.0)), float(np.linalg.norm(y - yhat) / (n2 if n2 > 0 else 1.0)), float(d.max(initial=0.0))
