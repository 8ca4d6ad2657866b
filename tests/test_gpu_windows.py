"""Rotation windows (several commuting rotations in one launch) against one
launch per rotation and against the CPU oracle.

A window replays the per-gate updates on each term's orbit with the same
operations as the one-rotation kernel, so every precision must give
bit-identical terms and identical work counters (updates, outs, peak) for every
window length.  FFX_WINDOW=1 switches windows off.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import pyffo
from _helpers import assert_same_terms
from fastfermion_amd import _abi, workloads as wl
from fastfermion_amd._native import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def run_windows(ctx, monkeypatch, fn):
    """fn() -> terms for window lengths 1..4, and for length 4 with a small
    launch grid (several chunks per block); returns [(terms, profile)]."""
    out = []
    for w, grid in ((1, ""), (2, ""), (3, ""), (4, ""), (4, "7")):
        monkeypatch.setenv("FFX_WINDOW", str(w))
        monkeypatch.setenv("FFX_WIN_GRID", grid or "0")
        t = fn()
        out.append((t, ctx.profile()))
    monkeypatch.delenv("FFX_WINDOW")
    monkeypatch.delenv("FFX_WIN_GRID")
    return out


def check_same(results, what=""):
    base, pb = results[0]
    for w, (t, p) in enumerate(results[1:], start=1):
        assert_same_terms(t, base, exact=True, what=f"{what} variant {w}")
        for key in ("updates", "outs", "peak_terms", "n_gates"):
            assert p[key] == pb[key], (what, w, key, p[key], pb[key])


def tfim(n, dt):
    g = [("ZZ", (i, (i + 1) % n), -2 * dt) for i in range(n)] + [("X", (i,), -2 * dt) for i in range(n)]
    return wl._pack(g)


PAULI_CASES = [
    ("tfim_batched", lambda: np.concatenate([tfim(8, 0.1)] * 5), 8, {"batched": True}),
    ("tfim_pergate_thresh", lambda: np.concatenate([tfim(8, 0.1)] * 5), 8, {"batched": False, "mincoeff": 1e-4}),
    ("tfim_maxdeg", lambda: np.concatenate([tfim(8, 0.1)] * 5), 8, {"maxdegree": 3}),
    ("tfim_xweight", lambda: np.concatenate([tfim(8, 0.1)] * 5), 8,
     {"batched": False, "max_xweight": 2, "xtrunc_period": 3}),
    ("kicked_6x6", lambda: wl.kicked_ising(6, 5)[0], 36, {"maxdegree": 8, "mincoeff": 1e-5}),
    ("kicked_6x6_pergate", lambda: wl.kicked_ising(6, 5)[0], 36,
     {"maxdegree": 8, "mincoeff": 1e-5, "batched": False}),
]


@pytest.mark.parametrize("name,gf,nq,kw", PAULI_CASES, ids=[c[0] for c in PAULI_CASES])
def test_pauli_windows_bit_exact(ctx, monkeypatch, name, gf, nq, kw):
    g = gf()
    W = _abi.words_for_bits(nq)
    x, z, c = wl.single_pauli("Z", [nq // 2], nq)
    res = run_windows(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, W, x, z, c, **kw))
    check_same(res, name)
    ref = pyffo.pauli_propagate(g, W, x, z, c, **kw)
    assert_same_terms(res[-1][0], ref, exact=True, what=name)
    assert res[-1][1]["updates"] == ref.updates


def test_heisenberg_dependent_axes(ctx, monkeypatch):
    """XX, YY, ZZ on one bond commute pairwise but ZZ = -XX*YY: a window may
    hold two of them only."""
    g = []
    for i in range(6):
        j = (i + 1) % 6
        g += [("XX", (i, j), 0.06), ("YY", (i, j), 0.06), ("ZZ", (i, j), 0.06)]
    g = np.concatenate([wl._pack(g)] * 4)
    x, z, c = wl.single_pauli("Z", [0], 6)
    res = run_windows(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, 1, x, z, c))
    check_same(res, "heisenberg")
    ref = pyffo.pauli_propagate(g, 1, x, z, c)
    assert_same_terms(res[-1][0], ref, exact=True)


def test_complex_polynomial_observable(ctx, monkeypatch):
    g = np.concatenate([tfim(7, 0.09)] * 4)
    x0, z0 = _abi.pauli_word("Z", [0], 1)
    x1, z1 = _abi.pauli_word("XY", [2, 3], 1)
    x = np.array([x0, x1], np.uint64).reshape(-1, 1)
    z = np.array([z0, z1], np.uint64).reshape(-1, 1)
    c = np.array([0.5 + 0.25j, -0.75j])
    kw = {"mincoeff": 1e-6, "batched": False}
    res = run_windows(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, 1, x, z, c, **kw))
    check_same(res, "complex")
    ref = pyffo.pauli_propagate(g, 1, x, z, c, **kw)
    assert_same_terms(res[-1][0], ref, exact=True)


@pytest.mark.parametrize("prec", [_abi.FP32, _abi.FP32C])
def test_fp32_windows_bit_identical(ctx, monkeypatch, prec):
    g, _ = wl.kicked_ising(5, 5, theta_lo=0.3, theta_hi=0.7)
    x, z, c = wl.single_pauli("Z", [12], 25)
    kw = dict(maxdegree=8, mincoeff=1e-7, batched=False, precision=prec)
    res = run_windows(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, 1, x, z, c, **kw))
    check_same(res, f"precision {prec}")


@pytest.mark.parametrize("seed", range(3))
def test_wide_keys_commuting_layers(ctx, monkeypatch, seed):
    """Layers of commuting random rotations on 130..250 qubits (3-4 words)."""
    rng = np.random.default_rng(300 + seed)
    nq = int(rng.integers(130, 250))
    W = _abi.words_for_bits(nq)
    recs = []
    for layer in range(6):
        ops = "Z" if layer % 2 == 0 else "X"
        for _ in range(12):
            w = int(rng.integers(1, 4))
            qs = sorted(rng.choice(nq, w, replace=False).tolist())
            recs.append((ops * w, tuple(qs), float(rng.uniform(-math.pi, math.pi))))
    g = wl._pack(recs)
    x, z, c = wl.single_pauli("Y", [int(rng.integers(0, nq))], nq)
    spec_q = [int(q) for q in rng.choice(nq, 3, replace=False)]
    xs, zs = [], []
    for q in spec_q:
        xa, za = _abi.pauli_word("X", [q], W)
        xs.append(xa)
        zs.append(za)
    x = np.concatenate([x, np.array(xs, np.uint64).reshape(-1, W)])
    z = np.concatenate([z, np.array(zs, np.uint64).reshape(-1, W)])
    c = np.array([1.0, 0.5, -0.25, 0.125], np.complex128)
    kw = {"maxdegree": 6, "mincoeff": 1e-7}
    res = run_windows(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, W, x, z, c, **kw))
    check_same(res, f"wide seed {seed}")
    ref = pyffo.pauli_propagate(g, W, x, z, c, **kw)
    assert_same_terms(res[-1][0], ref, exact=True)


def majorana_layers(rng, nm, layers):
    """Quadratic and quartic Majorana rotations in layers of disjoint supports
    (pairwise commuting inside a layer, overlapping across layers)."""
    gs = []
    for _ in range(layers):
        perm = rng.permutation(nm).tolist()
        while len(perm) >= 4:
            d = 2 if rng.random() < 0.6 else 4
            idx = sorted(perm[:d])
            perm = perm[d:]
            gs.append((idx, float(rng.uniform(-1.5, 1.5)), (d >> 1) & 1))
    return wl.pack_majorana(gs)


@pytest.mark.parametrize("nm", [16, 100, 200])
def test_majorana_windows_bit_exact(ctx, monkeypatch, nm):
    rng = np.random.default_rng(nm)
    W = _abi.words_for_bits(nm)
    g = majorana_layers(rng, nm, 4)
    m = np.zeros((2, W), np.uint64)
    m[0, 0] = 0b11
    m[1, 0] = 0b111100
    c = np.array([0.5j, 0.3])
    for maxdeg, minc in ((6, 1e-5), (8, 1e-3)):
        res = run_windows(ctx, monkeypatch,
                          lambda: ctx.majorana_propagate(g, W, m, c, maxdegree=maxdeg, mincoeff=minc))
        check_same(res, f"majorana nm={nm}")
        ref = pyffo.majorana_propagate(g, W, m, c, maxdegree=maxdeg, mincoeff=minc)
        assert_same_terms(res[-1][0], ref, exact=True, what=f"nm={nm} maxdeg={maxdeg}")


@pytest.mark.parametrize("grid", ["0", "1"])
@pytest.mark.parametrize("every", [1, 16])
def test_staged_appends_overflow_and_flush(ctx, monkeypatch, grid, every):
    """Appends of a window launch are staged in LDS per block and written out
    with one counter bump.  Terms X0 X1 X2 X3 (x) (distinct Z patterns on qubits
    4..14) have 15 absent orbit members each under Z rotations on qubits 0..3
    (one window of 4); the other terms carry Z0 and stay out of every orbit.
    every=1: each tile has ~3800 fresh members, more than the stage holds, and
    appends directly; every=16: ~240 per tile are staged, and a block that walks
    several tiles (a one-block grid) flushes its stage on the way."""
    n_terms = 300 if every == 1 else 1200
    xs, zs = [], []
    for p in range(1, n_terms + 1):
        bits = [b for b in range(11) if (p >> b) & 1]
        if p % every == 0:
            ops, qs = "XXXX" + "Z" * len(bits), [0, 1, 2, 3] + [4 + b for b in bits]
        else:
            ops, qs = "Z" + "Z" * len(bits), [0] + [4 + b for b in bits]
        x, z = _abi.pauli_word(ops, qs, 1)
        xs.append(x)
        zs.append(z)
    x = np.array(xs, np.uint64).reshape(-1, 1)
    z = np.array(zs, np.uint64).reshape(-1, 1)
    rng = np.random.default_rng(5)
    c = rng.normal(size=n_terms).astype(np.complex128)
    gates = wl._pack([("Z", (q,), float(rng.uniform(-1.2, 1.2))) for q in range(4)] * 3)
    kw = dict(maxdegree=16, mincoeff=0.0, batched=True)
    monkeypatch.setenv("FFX_WIN_GRID", grid)
    got = ctx.pauli_propagate(gates, 1, x, z, c, **kw)
    pg = ctx.profile()
    monkeypatch.setenv("FFX_WINDOW", "1")
    ref = ctx.pauli_propagate(gates, 1, x, z, c, **kw)
    pr = ctx.profile()
    monkeypatch.delenv("FFX_WINDOW")
    monkeypatch.delenv("FFX_WIN_GRID")
    active = n_terms // every
    assert len(got) == n_terms + 15 * active
    assert_same_terms(got, ref, exact=True, what=f"staged appends grid={grid} every={every}")
    for k in ("updates", "outs", "peak_terms"):
        assert pg[k] == pr[k], (k, pg[k], pr[k])
    orc = pyffo.pauli_propagate(gates, 1, x, z, c, **kw)
    assert_same_terms(got, orc, exact=True, what="staged appends vs oracle")
