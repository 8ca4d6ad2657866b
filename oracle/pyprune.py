"""CPU restatement of the learned term pruner -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker.  The pruner has no counterpart in the
reference (its truncation is degree, |c|, top-k and X/Y weight only,
src/pauli/truncate.h:16-57), so nothing here follows a reference file; it
restates what the device computes, from the mathematics:

* term_features: the 64 features of fastfermion_amd/csrc/ffx_prune.h
  (term_features), as exact bf16 values;
* mlp_forward / mlp_loss_grad: MLP 64 -> 128 -> 128 -> 1 with ReLU, in float64,
  rounding to bf16 exactly where the device does (inputs, weights, h1, dz2,
  dz1); with rounding=False it is the plain float64 network, whose analytic
  gradient tests/test_pruner_cpu.py checks against central finite differences.

Pinning: mlp_loss_grad by finite differences (exact-math check); bf16 by
exhaustive comparison with torch's bfloat16 conversion; term_features by
invariants (histogram totals equal the degree, hand-built strings).
"""
from __future__ import annotations

import math

import numpy as np

FEAT, HID = 64, 128
NPARAMS = FEAT * HID + HID + HID * HID + HID + HID + 1
OFF_W1, OFF_B1 = 0, FEAT * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    return (p[OFF_W1:OFF_B1].reshape(FEAT, HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def _r(a, rounding):
    return bf16(np.asarray(a, np.float32)).astype(np.float64) if rounding else np.asarray(a, np.float64)


def mlp_forward(p, X, rounding=True):
    """Predictions of the network on rows X (n x 64)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = _r(X, rounding)
    z1 = x @ _r(W1, rounding) + b1.astype(np.float64)
    h1 = _r(np.maximum(z1, 0.0), rounding)
    z2 = h1 @ _r(W2, rounding) + b2.astype(np.float64)
    h2 = np.maximum(z2, 0.0)
    return h2 @ _r(w3, rounding) + float(b3)


def mlp_loss_grad(p, X, T, rounding=True, under_weight=1.0):
    """Loss mean(w * (y - T)^2), w = under_weight where y < T else 1, and its gradient (flat)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = len(X)
    x = _r(X, rounding)
    W1r, W2r, w3r = _r(W1, rounding), _r(W2, rounding), _r(w3, rounding)
    z1 = x @ W1r + b1.astype(np.float64)
    h1 = _r(np.maximum(z1, 0.0), rounding)
    z2 = h1 @ W2r + b2.astype(np.float64)
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3r + float(b3)
    err = y - np.asarray(T, np.float64)
    wgt = np.where(err < 0, under_weight, 1.0)
    loss = float(np.mean(wgt * err * err))
    dy = 2.0 * wgt * err / n
    dz2 = (z2 > 0) * dy[:, None] * w3r[None, :]
    dz2r = _r(dz2, rounding)
    dh1 = dz2r @ W2r.T
    dz1 = (z1 > 0) * dh1
    dz1r = _r(dz1, rounding)
    g = np.zeros(NPARAMS)
    g[OFF_W1:OFF_B1] = (x.T @ dz1r).reshape(-1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2r).reshape(-1)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    return loss, g


def _bf(v):
    return float(bf16(np.float32(v)))


def term_features(key, coef_abs, kind, row, col, obs_site, rem_layers, mean_theta, frac_done, n_words):
    """Feature row of one term.  key: Pauli (x words, z words) or Majorana
    occupation words; kind 'pauli' or 'majorana'; row/col per qubit or
    Majorana index."""
    f = np.zeros(FEAT, np.float64)
    orow, ocol = row[obs_site], col[obs_site]
    sites = []
    n1 = n2 = n3 = 0
    diag = True
    if kind == "pauli":
        xw, zw = key
        for w in range(n_words):
            x, z = int(xw[w]), int(zw[w])
            n1 += bin(x & ~z & (2 ** 64 - 1)).count("1")
            n2 += bin(x & z).count("1")
            n3 += bin(z & ~x & (2 ** 64 - 1)).count("1")
            if x:
                diag = False
            b = x | z
            for t in range(64):
                if (b >> t) & 1:
                    sites.append((64 * w + t, bool((x >> t) & 1)))
    else:
        (mw,) = key
        for w in range(n_words):
            m = int(mw[w])
            pairs = m & (m >> 1) & 0x5555555555555555
            paired = pairs | (pairs << 1)
            n1 += bin(pairs).count("1")
            n2 += bin(m & ~paired & (2 ** 64 - 1)).count("1")
            if m & ~paired & (2 ** 64 - 1):
                diag = False
            for t in range(64):
                if (m >> t) & 1:
                    sites.append((64 * w + t, not ((paired >> t) & 1)))
    deg = len(sites)
    rs = [row[q] for q, _ in sites]
    cs = [col[q] for q, _ in sites]
    for (q, off) in sites:
        f[10 + min(row[q], 15)] += 0.25
        f[26 + min(col[q], 15)] += 0.25
        if off:
            f[42 + min(row[q], 15)] += 0.25
    f[0] = deg * 0.125
    f[1] = n1 * 0.125
    f[2] = n2 * 0.125
    f[3] = n3 * 0.125
    f[4] = rem_layers / 32.0
    f[5] = mean_theta
    logc = math.log10(coef_abs) if coef_abs > 0 else -300.0
    f[6] = max(logc, -40.0) * 0.125
    f[7] = frac_done
    if deg:
        f[8] = (max(rs) - min(rs)) / 16.0
        f[9] = (max(cs) - min(cs)) / 16.0
        f[58] = sum(rs) / deg / 16.0
        f[59] = sum(cs) / deg / 16.0
        f[61] = min(abs(r - orow) + abs(c - ocol) for r, c in zip(rs, cs)) / 16.0
    f[60] = 1.0 if diag else 0.0
    f[63] = 1.0
    return bf16(f.astype(np.float32))
