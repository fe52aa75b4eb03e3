"""Python mirror of entropical's density call surface over the HIP C ABI.

Names, argument meaning and error behaviour follow the reference:
  combined_densities(coords, n_rows, i_cols, h[, tau])  densities_cuda.hpp:10-24
  epa_convolution(sorted_coords, h)                     densities_cuda.hpp:29-31
  prob_dens_coord_prep(coords, n_rows, i_cols, h, tau, row_major)
                                                        densities_common.hpp:11-18
Coordinates are column-major: ``coords[col * n_rows + row]`` (a flat array or
an ``(n_cols, n_rows)`` C-contiguous array).  Errors raise ``EntropicalError``
where the reference prints to stderr and exits.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import EntropicalError, check, load

_METHODS = {"cell": _lib.METHOD_CELL, "brute": _lib.METHOD_BRUTE}


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    n = C.c_int(0)
    check(load().ent_device_count(C.byref(n)))
    return n.value


class Context:
    """A device, its stream and the engine workspace (ent_ctx)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        check(lib.ent_ctx_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise EntropicalError(_lib.ENT_ERR_ARG, "context closed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            load().ent_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tuning(self, q: int = 2, m0: int = 8, mr: int = 1, group_h: int = 4, sub: int = 0) -> None:
        check(load().ent_ctx_set_tuning(self.handle, q, m0, mr, group_h, sub))

    def get_tuning(self) -> tuple:
        """(q, m0, mr, group_h, sub)."""
        t = (C.c_int * 5)()
        check(load().ent_ctx_get_tuning(self.handle, t))
        return tuple(int(v) for v in t)

    def set_graph(self, enable: bool) -> None:
        check(load().ent_ctx_set_graph(self.handle, 1 if enable else 0))

    def last_stats(self) -> dict:
        dm, tm = C.c_double(), C.c_double()
        pairs, tiles = C.c_uint64(), C.c_uint64()
        check(load().ent_ctx_last_stats(self.handle, C.byref(dm), C.byref(tm), C.byref(pairs), C.byref(tiles)))
        by_d = (C.c_uint64 * 4)()
        check(load().ent_ctx_last_pairs(self.handle, by_d))
        srt = (C.c_uint32 * 2)()
        check(load().ent_ctx_last_sort(self.handle, srt))
        mom = (C.c_uint64 * 4)()
        check(load().ent_ctx_last_moments(self.handle, mom))
        tl = (C.c_uint64 * 2)()
        check(load().ent_ctx_last_tiles(self.handle, tl))
        gr = C.c_uint32()
        check(load().ent_ctx_last_graph_replays(self.handle, C.byref(gr)))
        return {"density_ms": dm.value, "total_ms": tm.value, "pairs": pairs.value, "tiles": tiles.value,
                "pairs_by_d": [int(v) for v in by_d[:3]], "lane_pairs": int(by_d[3]),
                "radix_passes": int(srt[0]), "sort_bases": int(srt[1]),
                "moment_pairs": int(mom[0]) + int(mom[1]), "moment_pairs_by_d": [int(mom[0]), int(mom[1])],
                "moment_terms": int(mom[2]) + int(mom[3]), "moment_terms_by_d": [int(mom[2]), int(mom[3])],
                "raw_tiles": int(tl[1]), "graph_replays": int(gr.value)}

    # device memory -------------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(load().ent_dev_alloc(self.handle, int(nbytes), C.byref(p)))
        return p.value

    def free(self, ptr: int) -> None:
        check(load().ent_dev_free(self.handle, C.c_void_p(ptr)))

    def h2d(self, dst: int, src: np.ndarray) -> None:
        src = np.ascontiguousarray(src)
        check(load().ent_h2d(self.handle, C.c_void_p(dst), _ptr(src), src.nbytes))

    def d2h(self, dst: np.ndarray, src: int) -> None:
        assert dst.flags["C_CONTIGUOUS"]
        check(load().ent_d2h(self.handle, _ptr(dst), C.c_void_p(src), dst.nbytes))

    def sync(self) -> None:
        check(load().ent_ctx_sync(self.handle))


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _as_colmajor(coords) -> np.ndarray:
    a = np.ascontiguousarray(coords, dtype=np.float32)
    return a.reshape(-1)


def _u32(v: Sequence[int]) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.uint32))


def _f32(v: Sequence[float]) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32))


def _check_dims(i_cols, h, tau) -> int:
    n_dim = len(i_cols)
    if n_dim < 1 or n_dim > 3:
        raise EntropicalError(_lib.ENT_ERR_ARG, "error: can only compute combined probabilities in 1, 2 or 3 dimensions.")
    if len(h) != n_dim:
        raise EntropicalError(_lib.ENT_ERR_ARG, "error: number of bandwidth parameters does not match number of selected columns.")
    if tau is not None and len(tau) != n_dim:
        raise EntropicalError(_lib.ENT_ERR_ARG, "error: number of lagtimes (tau) does not match number of selected dimensions.")
    return n_dim


def combined_densities(coords, n_rows: int, i_cols: Sequence[int], h: Sequence[float],
                       tau: Optional[Sequence[int]] = None, method: str = "cell",
                       ctx: Optional[Context] = None) -> np.ndarray:
    """Densities of the selected (lagged) columns at every selected row.

    Returns float64 of length ``n_rows - max(tau)``, each value divided by
    the original ``n_rows`` (densities_cuda.cu:46-105).
    """
    n_dim = _check_dims(i_cols, h, tau)
    flat = _as_colmajor(coords)
    if n_rows <= 0 or flat.size % n_rows:
        raise EntropicalError(_lib.ENT_ERR_ARG, "coords size is not a multiple of n_rows")
    n_cols = flat.size // n_rows
    ic, hh = _u32(i_cols), _f32(h)
    tt = _u32(tau) if tau is not None else None
    tau_max = int(tt.max()) if tt is not None else 0
    out = np.empty(max(n_rows - tau_max, 0), dtype=np.float64)
    ctx = ctx or default_context()
    check(load().ent_combined_densities(
        ctx.handle, _ptr(flat), n_rows, n_cols, _ptr(ic), _ptr(hh),
        _ptr(tt) if tt is not None else None, n_dim, _METHODS[method], _ptr(out)))
    return out


def prob_dens_coord_prep(coords, n_rows: int, i_cols: Sequence[int], h: Sequence[float],
                         tau: Sequence[int], row_major_result: bool) -> np.ndarray:
    """Lagged column gather (densities_common.hxx:7-50), flat float32."""
    n_dim = _check_dims(i_cols, h, tau)
    flat = _as_colmajor(coords)
    ic, hh, tt = _u32(i_cols), _f32(h), _u32(tau)
    ns = n_rows - int(tt.max())
    out = np.empty(max(ns, 0) * n_dim, dtype=np.float32)
    n_sel = C.c_size_t()
    check(load().ent_coord_prep(_ptr(flat), n_rows, _ptr(ic), _ptr(hh), _ptr(tt), n_dim,
                                1 if row_major_result else 0, _ptr(out), C.byref(n_sel)))
    return out


def epa_convolution(sorted_coords, h: float, ctx: Optional[Context] = None) -> float:
    """(1/N^2) sum_i sum_j [d<=2](4.5 - 2.25 d), d=|x_i-x_j|/h (densities_omp.cpp:235-277)."""
    x = np.ascontiguousarray(sorted_coords, dtype=np.float32)
    out = C.c_float()
    ctx = ctx or default_context()
    check(load().ent_epa_convolution(ctx.handle, _ptr(x), x.size, float(h), C.byref(out)))
    return out.value


def bin_column(x, box_size: float, ctx: Optional[Context] = None):
    """GPU Boxes binning of one column (boxedsearch.cpp:15-34, with the total fixes).

    Returns (box ids int32, n_boxes, min, max).
    """
    xx = np.ascontiguousarray(x, dtype=np.float32)
    box = np.empty(xx.size, dtype=np.int32)
    nb, mn, mx = C.c_uint32(), C.c_float(), C.c_float()
    ctx = ctx or default_context()
    check(load().ent_bin(ctx.handle, _ptr(xx), xx.size, float(box_size), _ptr(box),
                         C.byref(nb), C.byref(mn), C.byref(mx)))
    return box, nb.value, mn.value, mx.value


def cell_sort(coords, n_rows: int, i_cols, h, tau=None, ctx: Optional[Context] = None):
    """Stable GPU cell sort of one job: (sorted keys u64, perm u32, nb, mn, bs)."""
    n_dim = _check_dims(i_cols, h, tau)
    flat = _as_colmajor(coords)
    n_cols = flat.size // n_rows
    ic, hh = _u32(i_cols), _f32(h)
    tt = _u32(tau) if tau is not None else None
    ns = n_rows - (int(tt.max()) if tt is not None else 0)
    keys = np.empty(ns, dtype=np.uint64)
    perm = np.empty(ns, dtype=np.uint32)
    nb = np.zeros(n_dim, dtype=np.uint32)
    mn = np.zeros(n_dim, dtype=np.float32)
    bs = np.zeros(n_dim, dtype=np.float32)
    ctx = ctx or default_context()
    check(load().ent_cell_sort(ctx.handle, _ptr(flat), n_rows, n_cols, _ptr(ic), _ptr(hh),
                               _ptr(tt) if tt is not None else None, n_dim, _ptr(keys), _ptr(perm),
                               _ptr(nb), _ptr(mn), _ptr(bs)))
    return keys, perm, nb, mn, bs


def combined_densities_dev(ctx: Context, x_dev: int, n_rows: int, n_cols: int, i_cols: Sequence[int],
                           h: Sequence[float], out_dev: int, tau: Optional[Sequence[int]] = None,
                           method: str = "cell", noalloc: bool = False) -> None:
    """combined_densities over device-resident column-major coordinates into
    a device array of n_rows - max(tau) doubles (ent_combined_densities_dev).
    noalloc: the variant that allocates nothing, its workspace being the
    device's fixed scratch array (ent_combined_densities_dev_noalloc; at most
    ENT_NOALLOC_MAX_POINTS points)."""
    n_dim = _check_dims(i_cols, h, tau)
    ic, hh = _u32(i_cols), _f32(h)
    tt = _u32(tau) if tau is not None else None
    fn = load().ent_combined_densities_dev_noalloc if noalloc else load().ent_combined_densities_dev
    check(fn(ctx.handle, C.c_void_p(x_dev), n_rows, n_cols, _ptr(ic), _ptr(hh),
             _ptr(tt) if tt is not None else None, n_dim, _METHODS[method], C.c_void_p(out_dev)))


def batch_densities_dev(ctx: Context, x_dev: int, jobs: Sequence[dict], dens_dev: int,
                        method: str = "cell") -> None:
    """Run many density jobs over device-resident coordinates (see ent_job)."""
    arr = (_lib.ent_job * len(jobs))()
    for a, j in zip(arr, jobs):
        a.n_dim = j["n_dim"]
        a.n = j["n"]
        a.n_rows = j["n_rows"]
        for e in range(3):
            a.off[e] = int(j["off"][e]) if e < len(j["off"]) else 0
            a.h[e] = float(j["h"][e]) if e < len(j["h"]) else 1.0
        a.out_off = int(j["out_off"])
    check(load().ent_batch_densities_dev(ctx.handle, C.c_void_p(x_dev), arr, len(jobs),
                                         _METHODS[method], C.c_void_p(dens_dev)))


class DeviceArray:
    """A float32 array uploaded once to device memory (freed with .free())."""

    def __init__(self, ctx: Context, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float32)
        self.ctx = ctx
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = ctx.alloc(host.nbytes)
        ctx.h2d(self.ptr, host)

    def free(self) -> None:
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = 0


def last_driver_ms(ctx: Context) -> float:
    ms = C.c_double()
    check(load().ent_ctx_last_driver_ms(ctx.handle, C.byref(ms)))
    return ms.value


def mi_matrix(ctx: Context, x: DeviceArray, h: Sequence[float], bits: bool = False,
              method: str = "cell") -> np.ndarray:
    """Mutual-information matrix of all column pairs (mi.cpp:14-69), one batch."""
    n_cols, n_rows = x.shape
    hh = _f32(h)
    if hh.size != n_cols:
        raise EntropicalError(_lib.ENT_ERR_ARG, "error: number of given bandwidth values does not match the number of selected columns!")
    out = np.empty((n_cols, n_cols), dtype=np.float64)
    check(load().ent_mi_matrix_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, _ptr(hh),
                                   int(bits), _METHODS[method], _ptr(out)))
    return out


def te_matrix(ctx: Context, x: DeviceArray, h: Sequence[float], taus: Sequence[int],
              bits: bool = False, method: str = "cell") -> np.ndarray:
    """Transfer-entropy matrices for each lag (transs.cpp:16-125), one batch.

    Returns (n_taus, n_cols, n_cols); [t, i, j] holds the reference's
    transs[i*n+j] for lag taus[t].
    """
    n_cols, n_rows = x.shape
    hh, tt = _f32(h), _u32(taus)
    if hh.size != n_cols:
        raise EntropicalError(_lib.ENT_ERR_ARG, "error: number of given bandwidth values does not match the number of selected columns!")
    out = np.empty((tt.size, n_cols, n_cols), dtype=np.float64)
    check(load().ent_te_matrix_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, _ptr(hh), _ptr(tt),
                                   tt.size, int(bits), _METHODS[method], _ptr(out)))
    return out


def read_xtc(path) -> np.ndarray:
    """XTC trajectory as a column-major (3*natoms, n_frames) float32 array:
    column 3a + e holds coordinate e of atom a (coords_file.cpp:95-114)."""
    nf, na = C.c_size_t(), C.c_uint32()
    p = str(path).encode()
    check(load().ent_read_xtc(p, None, 0, C.byref(nf), C.byref(na)))
    out = np.empty((3 * na.value, nf.value), dtype=np.float32)
    check(load().ent_read_xtc(p, _ptr(out), out.size, C.byref(nf), C.byref(na)))
    return out


def _hcheck(x: DeviceArray, h: Sequence[float]) -> np.ndarray:
    hh = _f32(h)
    if hh.size != x.shape[0]:
        raise EntropicalError(_lib.ENT_ERR_ARG, "error: number of given bandwidth values does not match the number of selected columns!")
    return hh


def autoent(ctx: Context, x: DeviceArray, h: Sequence[float], tau: int, bits: bool = False,
            method: str = "cell") -> np.ndarray:
    """AutoEnt::main (autoent.cpp:16-63) for every column, one batch."""
    hh = _hcheck(x, h)
    n_cols, n_rows = x.shape
    out = np.empty(n_cols, dtype=np.float64)
    check(load().ent_autoent_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, _ptr(hh), int(tau), int(bits),
                                 _METHODS[method], _ptr(out)))
    return out


def kldiv(ctx: Context, x: DeviceArray, h: Sequence[float], bits: bool = False, method: str = "cell") -> np.ndarray:
    """KLDiv::main (kldiv.cpp:14-63): (n_cols, n_cols) matrix, one batch."""
    hh = _hcheck(x, h)
    n_cols, n_rows = x.shape
    out = np.empty((n_cols, n_cols), dtype=np.float64)
    check(load().ent_kldiv_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, _ptr(hh), int(bits),
                               _METHODS[method], _ptr(out)))
    return out


def negs(ctx: Context, x: DeviceArray, h: Sequence[float], bits: bool = False, method: str = "cell") -> np.ndarray:
    """Negs::main (negs.cpp:14-53): per-column value, one batch."""
    hh = _hcheck(x, h)
    n_cols, n_rows = x.shape
    out = np.empty(n_cols, dtype=np.float64)
    check(load().ent_negs_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, _ptr(hh), int(bits),
                              _METHODS[method], _ptr(out)))
    return out


def hestimate(ctx: Context, x: DeviceArray, kind: str = "amise") -> np.ndarray:
    """Bandwidth per column (Hestimate::main, hestimate.hxx:6-24): "thumb"
    (hestimate.cpp:19-30) or "amise" (hestimate.cpp:41-104, all columns'
    interval searches batched on the device)."""
    kinds = {"thumb": _lib.HEST_THUMB, "amise": _lib.HEST_AMISE}
    if kind not in kinds:
        raise ValueError(f"unknown bandwidth estimator {kind!r}")
    n_cols, n_rows = x.shape
    out = np.empty(n_cols, dtype=np.float32)
    check(load().ent_hestimate_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, kinds[kind], _ptr(out)))
    return out


def make_jobs(specs: Sequence[dict]):
    arr = (_lib.ent_job * len(specs))()
    for a, j in zip(arr, specs):
        a.n_dim = j["n_dim"]
        a.n = j["n"]
        a.n_rows = j.get("n_rows", j["n"])
        for e in range(3):
            a.off[e] = int(j["off"][e]) if e < len(j["off"]) else 0
            a.h[e] = float(j["h"][e]) if e < len(j["h"]) else 1.0
        a.out_off = int(j.get("out_off", 0))
    return arr


def batch_entropy(ctx: Context, x_dev_ptr: int, specs: Sequence[dict], method: str = "cell") -> np.ndarray:
    """H = -mean ln p for every job of a batch (jobs index device memory)."""
    arr = make_jobs(specs)
    out = np.empty(len(specs), dtype=np.float64)
    check(load().ent_batch_entropy_dev(ctx.handle, C.c_void_p(x_dev_ptr), arr, len(specs),
                                       _METHODS[method], _ptr(out)))
    return out


def mi_autotune(ctx: Context, x: DeviceArray, h: Sequence[float]) -> tuple:
    """Tune the engine for the MI job set of ``x`` (kept on ``ctx``)."""
    n_cols, n_rows = x.shape
    hh = _f32(h)
    t = (C.c_int * 5)()
    check(load().ent_mi_autotune_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, _ptr(hh), t))
    return tuple(int(v) for v in t)


def te_autotune(ctx: Context, x: DeviceArray, h: Sequence[float], taus: Sequence[int]) -> tuple:
    """Tune the engine for the transfer-entropy job set (kept on ``ctx``)."""
    n_cols, n_rows = x.shape
    hh, tt = _f32(h), _u32(taus)
    t = (C.c_int * 5)()
    check(load().ent_te_autotune_dev(ctx.handle, C.c_void_p(x.ptr), n_rows, n_cols, _ptr(hh), _ptr(tt),
                                     tt.size, t))
    return tuple(int(v) for v in t)


def batch_autotune(ctx: Context, x_dev_ptr: int, specs: Sequence[dict], return_ms: bool = False):
    """Tune the engine for an arbitrary batch of jobs (kept on ``ctx``).

    Returns the tuning (q, m0, mr, group_h, sub); with ``return_ms`` also the
    winner's time on the whole batch in ms (-1 when the tuning came from the
    cache of an earlier search of the same workload signature)."""
    arr = make_jobs(specs)
    t = (C.c_int * 5)()
    ms = C.c_double()
    check(load().ent_autotune_dev(ctx.handle, C.c_void_p(x_dev_ptr), arr, len(specs), t, C.byref(ms)))
    tun = tuple(int(v) for v in t)
    return (tun, ms.value) if return_ms else tun
