"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NP_KEYS = ("eps", "rs", "rs_seg", "analytic_calc", "n_start", "n_start_seg", "p_max", "p_max_seg")


_CACHE = {}


def golden(name):
    """All arrays of a fixture, decompressed once."""
    if name not in _CACHE:
        with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as f:
            _CACHE[name] = {k: f[k] for k in f.files}
    return _CACHE[name]


def np_dict(row):
    d = dict(zip(NP_KEYS, [float(v) for v in row]))
    for kk in ("n_start", "n_start_seg", "p_max", "p_max_seg"):
        d[kk] = int(d[kk])
    return d


def rel_err(a, b, floor=0.0):
    """max |a - b| / max(max |b|, floor); non-finite entries must agree exactly
    (NaN where the reference has NaN, the same infinity where it has one)."""
    a, b = np.asarray(a), np.asarray(b)
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a = np.asarray(a, dtype=np.complex128)
        b = np.asarray(b, dtype=np.complex128)
        a = np.stack([a.real, a.imag], axis=-1)
        b = np.stack([b.real, b.imag], axis=-1)
    na, nb = ~np.isfinite(a), ~np.isfinite(b)
    assert np.array_equal(na, nb), "non-finite pattern differs"
    assert np.array_equal(np.isnan(a[na]), np.isnan(b[nb])), "NaN/inf differ"
    assert np.array_equal(a[na & ~np.isnan(a)], b[nb & ~np.isnan(b)]), "infinities differ"
    a = np.where(na, 0, a)
    b = np.where(nb, 0, b)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, floor)
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b))) / scale


def rowwise_rel(a, b):
    """Per row: |a_i - b_i| / |b_i| (vectors along the last axis), NaN rows must agree."""
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    na, nb = np.isnan(a).any(axis=1), np.isnan(b).any(axis=1)
    assert np.array_equal(na, nb), "NaN pattern differs"
    a, b = a[~na], b[~nb]
    den = np.linalg.norm(b, axis=1)
    num = np.linalg.norm(a - b, axis=1)
    den = np.where(den == 0, 1.0, den)
    return num / den


def block_metric(Za, Zb):
    """Per 2x2 block ||dB||_F / max(||B_ref||_F, 1e-3 * median block norm)."""
    n2 = Zb.shape[0]
    A = Za.reshape(n2 // 2, 2, -1, 2)
    B = Zb.reshape(n2 // 2, 2, -1, 2)
    nb = np.sqrt((np.abs(B) ** 2).sum(axis=(1, 3)))
    nd = np.sqrt((np.abs(A - B) ** 2).sum(axis=(1, 3)))
    floor = 1e-3 * np.median(nb)
    return nd / np.maximum(nb, floor)
