"""ctypes binding of the CPU oracle (oracle/libkde_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed host
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libkde_oracle.so"
REF_LIB = HERE / "_ref" / "libref_harness.so"

_lib = None
_ref = None

P = C.c_void_p
SZ = C.c_size_t


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle not built: {LIB} (make -C oracle)")
        L = C.CDLL(str(LIB))
        sig = {
            "orc_set_threads": (None, [C.c_int]),
            "orc_max_threads": (C.c_int, []),
            "orc_coord_prep": (C.c_long, [P, SZ, P, P, C.c_uint, C.c_int, P]),
            "orc_density_fp64": (C.c_long, [P, SZ, P, P, P, C.c_uint, P]),
            "orc_density_fp64_window": (C.c_long, [P, SZ, P, P, P, C.c_uint, P]),
            "orc_density_cuda_replica": (C.c_long, [P, SZ, P, P, P, C.c_uint, P]),
            "orc_density_omp_packed": (C.c_long, [P, SZ, C.c_uint, P, P]),
            "orc_bin": (C.c_int, [P, SZ, C.c_float, P, C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
            "orc_counting_sort": (C.c_int, [P, SZ, C.c_uint32, P, P]),
            "orc_epa_conv_omp": (C.c_float, [P, SZ, C.c_float]),
            "orc_epa_conv_fp64": (C.c_double, [P, SZ, C.c_float]),
            "orc_thumb": (C.c_float, [P, SZ]),
            "orc_amise": (C.c_float, [P, SZ, C.c_int]),
            "orc_amise_value": (C.c_float, [C.c_float, C.c_float, SZ]),
            "orc_mi_ref": (C.c_float, [P, P, P, SZ, C.c_int]),
            "orc_mi_fp64": (C.c_double, [P, P, P, SZ, C.c_int]),
            "orc_te_ref": (C.c_double, [P, P, P, P, SZ, C.c_double, C.c_int, C.POINTER(C.c_long)]),
            "orc_te_fp64": (C.c_double, [P, P, P, P, SZ, C.c_double, C.c_int]),
            "orc_autoent_ref": (C.c_double, [P, P, SZ, C.c_float, C.c_int]),
            "orc_sum1_normalized": (None, [P, SZ, P]),
            "orc_kldiv_ref": (None, [P, P, SZ, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_negs_ref": (C.c_float, [P, SZ, C.c_int]),
            "orc_entropy": (C.c_double, [P, SZ]),
            "orc_loo_loglik": (None, [P, SZ, C.c_uint, P, P, C.c_uint, C.c_double, P]),
            "orc_count_in_support": (C.c_uint64, [P, SZ, C.c_uint, P]),
            "orc_count_in_support_jobs": (C.c_uint64, [P, SZ, C.c_uint, C.c_uint, P, P]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def ref_available() -> bool:
    return REF_LIB.exists()


def ref() -> C.CDLL:
    """Harness over the reference's own sources (built only where they exist)."""
    global _ref
    if _ref is None:
        L = C.CDLL(str(REF_LIB))
        L.ref_coord_prep.restype = C.c_long
        L.ref_coord_prep.argtypes = [P, SZ, P, P, P, C.c_uint, C.c_int, P]
        L.ref_boxes.restype = C.c_long
        L.ref_boxes.argtypes = [P, SZ, SZ, C.c_float, P]
        L.ref_neighbors.restype = C.c_long
        L.ref_neighbors.argtypes = [P, SZ, SZ, C.c_float, SZ, P, SZ]
        _ref = L
    return _ref


def _u32(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.uint32))


def _f32(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def _density(fn, coords, n_rows, i_cols, h, tau=None) -> np.ndarray:
    flat = _f32(coords).reshape(-1)
    d = len(i_cols)
    tau = [0] * d if tau is None else tau
    ic, hh, tt = _u32(i_cols), _f32(h), _u32(tau)
    out = np.empty(n_rows - int(tt.max()), dtype=np.float64)
    r = fn(_p(flat), n_rows, _p(ic), _p(hh), _p(tt), d, _p(out))
    if r < 0:
        raise ValueError(f"oracle density rejected input ({r})")
    return out


def density_fp64(coords, n_rows, i_cols, h, tau=None) -> np.ndarray:
    return _density(lib().orc_density_fp64, coords, n_rows, i_cols, h, tau)


def density_fp64_window(coords, n_rows, i_cols, h, tau=None) -> np.ndarray:
    return _density(lib().orc_density_fp64_window, coords, n_rows, i_cols, h, tau)


def density_cuda_replica(coords, n_rows, i_cols, h, tau=None) -> np.ndarray:
    return _density(lib().orc_density_cuda_replica, coords, n_rows, i_cols, h, tau)


def density_omp_packed(packed_colmajor, n, h) -> np.ndarray:
    x = _f32(packed_colmajor).reshape(-1)
    d = len(h)
    assert x.size == n * d
    hh = _f32(h)
    out = np.empty(n, dtype=np.float32)
    r = lib().orc_density_omp_packed(_p(x), n, d, _p(hh), _p(out))
    if r < 0:
        raise ValueError("oracle OMP density rejected input")
    return out


def coord_prep(coords, n_rows, i_cols, tau, row_major) -> np.ndarray:
    flat = _f32(coords).reshape(-1)
    ic, tt = _u32(i_cols), _u32(tau)
    d = len(i_cols)
    out = np.empty((n_rows - int(tt.max())) * d, dtype=np.float32)
    r = lib().orc_coord_prep(_p(flat), n_rows, _p(ic), _p(tt), d, 1 if row_major else 0, _p(out))
    if r < 0:
        raise ValueError("bad args")
    return out


def bin_column(x, box_size):
    xx = _f32(x)
    box = np.empty(xx.size, dtype=np.int32)
    nb, mn, mx = C.c_uint32(), C.c_float(), C.c_float()
    r = lib().orc_bin(_p(xx), xx.size, float(box_size), _p(box), C.byref(nb), C.byref(mn), C.byref(mx))
    if r != 0:
        raise ValueError("oracle bin rejected input")
    return box, nb.value, mn.value, mx.value


def counting_sort(keys, n_keys):
    k = _u32(keys)
    perm = np.empty(k.size, dtype=np.uint32)
    start = np.empty(int(n_keys) + 1, dtype=np.uint32)
    r = lib().orc_counting_sort(_p(k), k.size, int(n_keys), _p(perm), _p(start))
    if r != 0:
        raise ValueError("key out of range")
    return perm, start


def epa_conv_omp(sorted_x, h) -> float:
    x = _f32(sorted_x)
    return float(lib().orc_epa_conv_omp(_p(x), x.size, float(h)))


def epa_conv_fp64(x, h) -> float:
    x = _f32(x)
    return float(lib().orc_epa_conv_fp64(_p(x), x.size, float(h)))


def thumb(x) -> float:
    x = _f32(x)
    return float(lib().orc_thumb(_p(x), x.size))


def amise(x, exact=True) -> float:
    x = _f32(x)
    return float(lib().orc_amise(_p(x), x.size, 1 if exact else 0))


def amise_value(conv, h, n) -> float:
    return float(lib().orc_amise_value(float(conv), float(h), int(n)))


def _d(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def mi_ref(p_ij, p_i, p_j, bits=False) -> float:
    a, b, c = _d(p_ij), _d(p_i), _d(p_j)
    return float(lib().orc_mi_ref(_p(a), _p(b), _p(c), a.size, int(bits)))


def mi_fp64(p_ij, p_i, p_j, bits=False) -> float:
    a, b, c = _d(p_ij), _d(p_i), _d(p_j)
    return float(lib().orc_mi_fp64(_p(a), _p(b), _p(c), a.size, int(bits)))


def te_ref(p3, p1, p2, ptau, pref, bits=False):
    a, b, c, d = _d(p3), _d(p1), _d(p2), _d(ptau)
    n = C.c_long()
    v = lib().orc_te_ref(_p(a), _p(b), _p(c), _p(d), a.size, float(pref), int(bits), C.byref(n))
    return float(v), n.value


def te_fp64(p3, p1, p2, ptau, pref, bits=False) -> float:
    a, b, c, d = _d(p3), _d(p1), _d(p2), _d(ptau)
    return float(lib().orc_te_fp64(_p(a), _p(b), _p(c), _p(d), a.size, float(pref), int(bits)))


def autoent_ref(p1, p2, h, bits=False) -> float:
    a, b = _d(p1), _d(p2)
    return float(lib().orc_autoent_ref(_p(a), _p(b), a.size, float(h), int(bits)))


def sum1_normalized(p) -> np.ndarray:
    a = _d(p)
    out = np.empty_like(a)
    lib().orc_sum1_normalized(_p(a), a.size, _p(out))
    return out


def kldiv_ref(pi, pj, bits=False):
    a, b = _d(pi), _d(pj)
    x, y = C.c_double(), C.c_double()
    lib().orc_kldiv_ref(_p(a), _p(b), a.size, int(bits), C.byref(x), C.byref(y))
    return x.value, y.value


def negs_ref(p, bits=False) -> float:
    a = _d(p)
    return float(lib().orc_negs_ref(_p(a), a.size, int(bits)))


def entropy(p) -> float:
    a = _d(p)
    return float(lib().orc_entropy(_p(a), a.size))


def loo_loglik(pts_rowmajor, d, scale, hs, floor_p) -> np.ndarray:
    x = _f32(pts_rowmajor).reshape(-1)
    n = x.size // d
    sc, hh = _f32(scale), _f32(hs)
    out = np.empty(hh.size, dtype=np.float64)
    lib().orc_loo_loglik(_p(x), n, d, _p(sc), _p(hh), hh.size, float(floor_p), _p(out))
    return out


def count_in_support(coords, n_rows, i_cols, h) -> int:
    """Ordered pairs (i, j), the self pair included, with |x_ie - x_je| < h_e
    in every dim (d = 1, 2): the pairs whose Epanechnikov term is nonzero."""
    c = _f32(coords).reshape(-1, n_rows)
    packed = np.ascontiguousarray(c[list(i_cols)].reshape(-1))
    hh = _f32(h)
    return int(lib().orc_count_in_support(_p(packed), n_rows, len(i_cols), _p(hh)))


def count_in_support_jobs(coords, n_rows, cols, h) -> int:
    """count_in_support summed over jobs of one dimension d (cols: n_jobs x d
    column indices, h: n_jobs x d bandwidths), OpenMP over the jobs."""
    c = _f32(coords).reshape(-1)
    cc = np.ascontiguousarray(np.asarray(cols, dtype=np.uint32))
    hh = np.ascontiguousarray(np.asarray(h, dtype=np.float32))
    n_jobs, d = cc.shape
    assert hh.shape == cc.shape
    return int(lib().orc_count_in_support_jobs(_p(c), n_rows, n_jobs, d, _p(cc), _p(hh)))


# ---- reference harness (present only where the reference tree was built)
def ref_coord_prep(coords, n_rows, i_cols, h, tau, row_major) -> np.ndarray:
    flat = _f32(coords).reshape(-1)
    ic, hh, tt = _u32(i_cols), _f32(h), _u32(tau)
    d = len(i_cols)
    out = np.empty((n_rows - int(tt.max())) * d, dtype=np.float32)
    ref().ref_coord_prep(_p(flat), n_rows, _p(ic), _p(hh), _p(tt), d, 1 if row_major else 0, _p(out))
    return out


def ref_boxes(coords, n_rows, col, box_size):
    flat = _f32(coords).reshape(-1)
    box = np.empty(n_rows, dtype=np.int64)
    nb = ref().ref_boxes(_p(flat), n_rows, col, float(box_size), _p(box))
    return box, int(nb)


# ------------------------------------------------ reference XTC reader/writer
# The reference's vendored xdrfile library (coords_file/xdrfile), compiled
# into _ref/libref_xdr.so by `make -C oracle ref` where the reference tree is
# present.
REF_XDR = HERE / "_ref" / "libref_xdr.so"
_xdr = None


def xdr_available() -> bool:
    return REF_XDR.exists()


def _xdr_lib() -> C.CDLL:
    global _xdr
    if _xdr is None:
        L = C.CDLL(str(REF_XDR))
        L.refx_write.restype = C.c_int
        L.refx_write.argtypes = [C.c_char_p, C.c_int, C.c_int, P, C.c_float]
        L.refx_read.restype = C.c_int
        L.refx_read.argtypes = [C.c_char_p, C.c_int, P, C.POINTER(C.c_int)]
        _xdr = L
    return _xdr


def ref_write_xtc(path, xyz, precision=1000.0) -> None:
    """xyz: (n_frames, natoms, 3) float32, written with the reference writer."""
    a = np.ascontiguousarray(xyz, dtype=np.float32)
    rc = _xdr_lib().refx_write(str(path).encode(), a.shape[1], a.shape[0], _p(a), float(precision))
    if rc != 0:
        raise RuntimeError(f"reference XTC writer failed ({rc})")


def ref_read_xtc(path, max_frames: int, natoms: int) -> np.ndarray:
    """(n_frames, natoms, 3) float32 read with the reference reader."""
    out = np.empty((max_frames, natoms, 3), dtype=np.float32)
    nf = C.c_int()
    na = _xdr_lib().refx_read(str(path).encode(), max_frames, _p(out), C.byref(nf))
    if na != natoms:
        raise RuntimeError(f"reference XTC reader returned {na} atoms")
    return out[: nf.value]
