"""Sparse-wavelength combo scan (include/sky_scan.h): the evaluation loop of
the reference's wavelength optimizer (crates/sky-wavelength-optimizer/src/
main.rs) on the GPU, plus its host-side pieces (band widths, quadrature
weights, CIE table, combo enumeration, screening indices).

    ds = ScanDataset(values, reference_oklab)       # [n, bands], [n, 3]
    s, sq, st = ds.scan(combos, coef, samples)      # screening sums
    stats = ds.refine(combos, coef, samples)        # mean, rmse, p50..max
"""
from __future__ import annotations

import ctypes as C
import itertools
from typing import Optional

import numpy as np

from . import _lib as L
from .api import check

VORONOI, CMF_MASS, CMF_MOMENTS = 0, 1, 2
STRATEGIES = {"voronoi": VORONOI, "cmf-mass": CMF_MASS, "cmf-moments": CMF_MOMENTS}
MAX_PRIMARIES = 8
EXHAUSTIVE_COMBO_LIMIT = 5_000_000  # main.rs:41

# Oklab (Ottosson 2020): XYZ -> LMS and LMS^(1/3) -> Lab (main.rs:1088-1098)
OKLAB_M1 = np.array([[0.8189330101, 0.3618667424, -0.1288597137],
                     [0.0329845436, 0.9293118715, 0.0361456387],
                     [0.0482003018, 0.2643662691, 0.6338517070]])
OKLAB_M2 = np.array([[0.2104542553, 0.7936177850, -0.0040720468],
                     [1.9779984951, -2.4285922050, 0.4505937099],
                     [0.0259040371, 0.7827717662, -0.8086757660]])

STAT_FIELDS = ("mean", "rmse", "p50", "p90", "p95", "p99", "max")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cie_1931_2deg(lambda_nm) -> np.ndarray:
    """[n, 3] CMF values (spectrum.rs:196-219)."""
    lam = np.ascontiguousarray(lambda_nm, dtype=np.float32).ravel()
    out = np.zeros((lam.size, 3), np.float64)
    check(L.lib().sky_cie_1931_2deg(_fp(lam), lam.size, _dp(out)))
    return out


def band_widths(wavelengths_nm) -> np.ndarray:
    """Voronoi widths of the band grid (main.rs:714-733)."""
    wl = np.ascontiguousarray(wavelengths_nm, dtype=np.float32)
    out = np.zeros(wl.size, np.float64)
    check(L.lib().sky_band_widths(_fp(wl), wl.size, _dp(out)))
    return out


def quadrature_weights(wavelengths_nm, combos, strategy=VORONOI, cmf=None, threads=0) -> np.ndarray:
    """[n_combos, count] weights in nm (main.rs:735-1070)."""
    wl = np.ascontiguousarray(wavelengths_nm, dtype=np.float32)
    cb = np.ascontiguousarray(np.atleast_2d(combos), dtype=np.uint8)
    if cmf is None and strategy != VORONOI:
        cmf = cie_1931_2deg(wl)
    cm = None if cmf is None else np.ascontiguousarray(cmf, dtype=np.float64)
    out = np.zeros(cb.shape, np.float64)
    check(L.lib().sky_quadrature_weights(_fp(wl), wl.size, None if cm is None else _dp(cm), cb.ctypes.data,
                                         cb.shape[1], cb.shape[0], strategy, threads, _dp(out)))
    return out


def combo_coefficients(combos, weights, widths, cmf, scale) -> np.ndarray:
    """c_k = weight_k / width_{b_k} * cmf(b_k) / scale, [n, count, 3] float32."""
    cb = np.atleast_2d(combos)
    c = (np.asarray(weights) / np.asarray(widths)[cb])[..., None] * np.asarray(cmf)[cb] / scale
    return np.ascontiguousarray(c, dtype=np.float32)


def xyz_to_oklab(xyz) -> np.ndarray:
    """FP64 Oklab of [..., 3] XYZ (main.rs:1088-1098)."""
    lms = np.asarray(xyz, dtype=np.float64) @ OKLAB_M1.T
    return np.cbrt(lms) @ OKLAB_M2.T


def all_combos(n_bands: int, count: int) -> np.ndarray:
    """Ascending combos in lexicographic order (main.rs:542-548, 693-712)."""
    return np.array(list(itertools.combinations(range(n_bands), count)), dtype=np.uint8).reshape(-1, count)


def screening_indices(sample_count: int, requested: int) -> np.ndarray:
    """Evenly spread sample subset (main.rs:1096-1105)."""
    if requested == 0 or requested >= sample_count:
        return np.arange(sample_count, dtype=np.uint32)
    stride = sample_count / requested
    idx = np.floor((np.arange(requested) + 0.5) * stride).astype(np.int64)
    return np.minimum(idx, sample_count - 1).astype(np.uint32)


class ScanDataset:
    """Band values and reference Oklab colours resident on one GPU."""

    def __init__(self, values: np.ndarray, reference_oklab: np.ndarray, device: int = 0):
        v = np.ascontiguousarray(values, dtype=np.float32)
        r = np.ascontiguousarray(reference_oklab, dtype=np.float32)
        assert v.ndim == 2 and r.shape == (v.shape[0], 3)
        self.n_samples, self.n_bands = v.shape
        h = C.c_void_p()
        check(L.lib().sky_scan_dataset_create(device, _fp(v), self.n_bands, self.n_samples, _fp(r), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().sky_scan_dataset_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _args(combos, coef, samples):
        cb = np.ascontiguousarray(np.atleast_2d(combos), dtype=np.uint8)
        cf = np.ascontiguousarray(coef, dtype=np.float32).reshape(cb.shape[0], cb.shape[1], 3)
        sm = None if samples is None else np.ascontiguousarray(samples, dtype=np.uint32)
        return cb, cf, sm

    def scan(self, combos, coef, samples: Optional[np.ndarray] = None):
        """Screening sums: (sum of errors, sum of squared errors, stats)."""
        cb, cf, sm = self._args(combos, coef, samples)
        s = np.zeros(cb.shape[0], np.float64)
        q = np.zeros(cb.shape[0], np.float64)
        st = L.ScanStats()
        check(L.lib().sky_scan_combos(self._h, cb.ctypes.data, _fp(cf), cb.shape[1], cb.shape[0],
                                      None if sm is None else sm.ctypes.data, 0 if sm is None else sm.size,
                                      _dp(s), _dp(q), C.byref(st)))
        return s, q, st

    def scan_device(self, combos, coef, out_sum, out_sq, samples=None, stream=None):
        """Screening sums with every operand a device tensor (uint8 [n, count],
        float32 [n, count, 3], float64 [n] outputs, optional uint32 samples),
        ordered on torch's current stream unless `stream` is given."""
        if stream is None:
            import torch
            stream = torch.cuda.current_stream(combos.device).cuda_stream
        n, count = combos.shape
        st = L.ScanStats()
        check(L.lib().sky_scan_combos_device(self._h, combos.data_ptr(), coef.data_ptr(), count, n,
                                             None if samples is None else samples.data_ptr(),
                                             0 if samples is None else samples.numel(), out_sum.data_ptr(),
                                             out_sq.data_ptr(), stream, C.byref(st)))
        return st

    def refine(self, combos, coef, samples: Optional[np.ndarray] = None):
        """Per combo mean, rmse, p50, p90, p95, p99, max (dict of arrays) and stats."""
        cb, cf, sm = self._args(combos, coef, samples)
        out = (L.ComboStats * cb.shape[0])()
        st = L.ScanStats()
        check(L.lib().sky_refine_combos(self._h, cb.ctypes.data, _fp(cf), cb.shape[1], cb.shape[0],
                                        None if sm is None else sm.ctypes.data, 0 if sm is None else sm.size,
                                        out, C.byref(st)))
        res = {f: np.array([getattr(o, f) for o in out]) for f in STAT_FIELDS}
        res["sample_count"] = np.array([o.sample_count for o in out])
        return res, st
