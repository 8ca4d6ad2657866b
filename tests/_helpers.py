"""Shared helpers of the test suite: canonical ordering and comparison of term sets."""
from __future__ import annotations

import numpy as np


def canon(x, z, c):
    """Sort terms by key; returns (keys, coeffs) with keys a 2-D uint64 array."""
    keys = x if z is None else np.concatenate([x, z], axis=1)
    if len(c) == 0:
        return keys.reshape(0, keys.shape[1] if keys.ndim == 2 else 0), np.zeros(0, np.complex128)
    order = np.lexsort(keys.T[::-1])
    return keys[order], np.asarray(c)[order]


def widen(keys, W_from, W_to, pauli=True):
    """Zero-extend term-major key words from W_from to W_to words per mask."""
    if W_from == W_to:
        return keys
    n = keys.shape[0]
    if pauli:
        out = np.zeros((n, 2 * W_to), np.uint64)
        out[:, :W_from] = keys[:, :W_from]
        out[:, W_to:W_to + W_from] = keys[:, W_from:]
    else:
        out = np.zeros((n, W_to), np.uint64)
        out[:, :W_from] = keys
    return out


def assert_same_terms(a, b, rtol=0.0, atol=0.0, exact=False, what=""):
    """a, b: objects with x, z (None for Majorana), c.  Keys must match exactly."""
    ka, ca = canon(a.x, a.z, a.c)
    kb, cb = canon(b.x, b.z, b.c)
    assert ka.shape == kb.shape, f"{what}: term count {ka.shape[0]} vs {kb.shape[0]}"
    assert np.array_equal(ka, kb), f"{what}: key sets differ"
    if exact:
        assert np.array_equal(ca.real, cb.real) and np.array_equal(ca.imag, cb.imag), \
            f"{what}: coefficients not bit-identical, max diff {np.max(np.abs(ca - cb)) if len(ca) else 0}"
    else:
        scale = np.maximum(np.abs(ca), np.abs(cb))
        bad = np.abs(ca - cb) > atol + rtol * scale
        assert not np.any(bad), f"{what}: {bad.sum()} coefficients differ, max {np.max(np.abs(ca - cb))}"


def rel_err(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def assert_close_terms(a, b, tol, mincoeff=0.0, what=""):
    """Approximate comparison for reduced-precision results: common keys agree to
    `tol` absolute; a key present on one side only must sit within `tol` of the
    truncation threshold (the only place a rounding difference can flip a key)."""
    ka, ca = canon(a.x, a.z, a.c)
    kb, cb = canon(b.x, b.z, b.c)
    da = {tuple(int(v) for v in ka[i]): ca[i] for i in range(len(ca))}
    db = {tuple(int(v) for v in kb[i]): cb[i] for i in range(len(cb))}
    worst = 0.0
    for k in set(da) | set(db):
        if k in da and k in db:
            worst = max(worst, abs(da[k] - db[k]))
        else:
            v = da.get(k, db.get(k))
            assert abs(v) <= mincoeff + tol, f"{what}: key only on one side with |c|={abs(v)}"
    assert worst <= tol, f"{what}: max coefficient difference {worst} > {tol}"
    return worst
