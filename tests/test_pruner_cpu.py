"""Pins of the pruner's CPU checker (oracle/pyprune.py), no GPU needed.

* bf16 rounding against torch's float32 -> bfloat16 conversion;
* the analytic gradient of the MLP (rounding off) against central finite
  differences in float64: a dropped term, a wrong sign or a transposed operand in
  the backward formulas fails this;
* term features on hand-built strings (counts, histograms, distance) and their
  invariants (histogram totals = degree / 4).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import pyprune
from fastfermion_amd import _abi, workloads as wl


def test_bf16_matches_torch():
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-3, 20000), rng.normal(0, 1e3, 20000),
                        np.array([0.0, -0.0, 1.0, -1.0, 0.25, 3.0e-39])]).astype(np.float32)
    ours = pyprune.bf16(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def small_params(rng):
    p = np.zeros(pyprune.NPARAMS, np.float64)
    p[pyprune.OFF_W1:pyprune.OFF_B1] = rng.normal(0, 0.3, 64 * 128)
    p[pyprune.OFF_B1:pyprune.OFF_W2] = rng.normal(0, 0.1, 128)
    p[pyprune.OFF_W2:pyprune.OFF_B2] = rng.normal(0, 0.15, 128 * 128)
    p[pyprune.OFF_B2:pyprune.OFF_W3] = rng.normal(0, 0.1, 128)
    p[pyprune.OFF_W3:pyprune.OFF_B3] = rng.normal(0, 0.2, 128)
    p[pyprune.OFF_B3] = -0.5
    return p


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = small_params(rng)
    X = rng.uniform(0, 1, (24, 64))
    T = rng.uniform(-2, 0, 24)

    def loss(q):
        # float64 forward without rounding (the parameters stay float64 here)
        W1, b1 = q[:8192].reshape(64, 128), q[8192:8320]
        W2, b2 = q[8320:24704].reshape(128, 128), q[24704:24832]
        w3, b3 = q[24832:24960], q[24960]
        h1 = np.maximum(X @ W1 + b1, 0)
        h2 = np.maximum(h1 @ W2 + b2, 0)
        y = h2 @ w3 + b3
        return np.mean((y - T) ** 2)

    # analytic gradient of the oracle with rounding off, evaluated in float64
    W1, b1, W2, b2, w3, b3 = (p[:8192].reshape(64, 128), p[8192:8320], p[8320:24704].reshape(128, 128),
                              p[24704:24832], p[24832:24960], p[24960])
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    l0 = np.mean((y - T) ** 2)
    _, g = pyprune.mlp_loss_grad(p.astype(np.float32), X, T, rounding=False)
    l32, _ = pyprune.mlp_loss_grad(p.astype(np.float32), X, T, rounding=False)
    assert abs(l32 - l0) < 1e-5 * max(1.0, l0)  # float32 parameters vs float64
    idx = list(rng.choice(8192, 6, replace=False)) + [8192 + 3, 8320 + 77, 8320 + 5000, 24704 + 9,
                                                        24832 + 11, 24832 + 100, 24960]
    q32 = p.astype(np.float32).astype(np.float64)
    for i in idx:
        h = 1e-5
        a, b = q32.copy(), q32.copy()
        a[i] += h
        b[i] -= h
        fd = (loss(a) - loss(b)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_features_hand_built_pauli():
    row, col = wl.lattice_geometry(4)
    W = 1
    x, z = _abi.pauli_word("XYZ", (0, 5, 14), W)  # sites (0,0), (1,1), (3,2)
    f = pyprune.term_features((x, z), 0.01, "pauli", row, col, obs_site=5, rem_layers=8, mean_theta=0.5,
                              frac_done=0.25, n_words=W)
    assert f[0] == 3 / 8 and f[1] == 1 / 8 and f[2] == 1 / 8 and f[3] == 1 / 8
    assert f[4] == 8 / 32 and f[5] == 0.5 and f[7] == 0.25
    assert f[6] == pyprune.bf16(np.float32(-2 * 0.125))
    assert f[8] == 3 / 16 and f[9] == 2 / 16
    assert f[10 + 0] == 0.25 and f[10 + 1] == 0.25 and f[10 + 3] == 0.25
    assert f[26 + 0] == 0.25 and f[26 + 1] == 0.25 and f[26 + 2] == 0.25
    assert f[42 + 0] == 0.25 and f[42 + 1] == 0.25 and f[42 + 3] == 0  # X at row 0, Y at row 1, Z off
    assert f[60] == 0 and f[61] == 0 and f[63] == 1
    assert abs(f[10:26].sum() - 3 * 0.25) < 1e-7 and abs(f[26:42].sum() - 3 * 0.25) < 1e-7


def test_features_hand_built_majorana():
    row, col = wl.hubbard_geometry(2, 2)
    m = np.zeros(1, np.uint64)
    for a in (0, 1, 6, 13):  # mode 0 paired (site 0), Majorana 6 (site 1), 13 (site 3)
        m[0] |= np.uint64(1) << np.uint64(a)
    f = pyprune.term_features((m,), 1.0, "majorana", row, col, obs_site=0, rem_layers=0, mean_theta=0.1,
                              frac_done=1.0, n_words=1)
    assert f[0] == 4 / 8 and f[1] == 1 / 8 and f[2] == 2 / 8 and f[3] == 0
    assert f[60] == 0 and f[61] == 0  # unpaired Majoranas: off-diagonal; site 0 is in the support
    assert abs(f[42:58].sum() - 2 * 0.25) < 1e-7


@pytest.mark.parametrize("seed", range(3))
def test_feature_histograms_total_degree(seed):
    rng = np.random.default_rng(seed)
    row, col = wl.lattice_geometry(10)
    x = np.zeros(2, np.uint64)
    z = np.zeros(2, np.uint64)
    qs = rng.choice(100, 7, replace=False)
    for q in qs:
        letter = "XYZ"[rng.integers(3)]
        w, b = divmod(int(q), 64)
        if letter in "XY":
            x[w] |= np.uint64(1) << np.uint64(b)
        if letter in "YZ":
            z[w] |= np.uint64(1) << np.uint64(b)
    f = pyprune.term_features((x, z), 0.5, "pauli", row, col, 55, 3, 0.4, 0.5, 2)
    assert f[0] == 7 / 8
    assert abs(f[10:26].sum() - 7 / 4) < 1e-6 and abs(f[26:42].sum() - 7 / 4) < 1e-6
    assert abs(f[42:58].sum() - (f[1] + f[2]) * 2) < 1e-6
