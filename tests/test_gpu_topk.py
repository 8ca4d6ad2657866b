"""Device top-k truncation (src/pauli/truncate.h:21-39) against the oracle.

Top-k keeps the k largest |c| after the threshold and before the X/Y cap.  The
reference erases surplus exact ties at the boundary in hash-map order, which
no other implementation can reproduce; the device keeps the ties with the
largest key hashes, the same set on every run and grid.  So the checks are:
the multiset of kept magnitudes is identical,
and every kept key strictly above the boundary magnitude is identical with a
bit-identical coefficient.  Random axes and angles make boundary ties (and
with them diverging later batches) unlikely; one case forces ties on purpose
and is checked at a single truncation point.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import pyffo
from _helpers import canon
from fastfermion_amd import _abi, workloads as wl
from fastfermion_amd._native import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def terms_dict(t):
    x, z, c = t.x, t.z, t.c
    return {tuple(int(v) for v in x[i]) + tuple(int(v) for v in z[i]): c[i] for i in range(len(c))}


def assert_topk_match(dev, ref, what=""):
    assert len(dev.c) == len(ref.c), (what, len(dev.c), len(ref.c))
    md = np.sort(np.abs(dev.c))
    mr = np.sort(np.abs(ref.c))
    assert np.array_equal(md, mr), what
    if len(mr) == 0:
        return
    edge = mr[0]
    dd, rd = terms_dict(dev), terms_dict(ref)
    for k, v in rd.items():
        if abs(v) > edge:
            assert k in dd and dd[k] == v, (what, k)


def random_rotations(rng, nq, ngates, maxw=3):
    recs = []
    for _ in range(ngates):
        w = int(rng.integers(1, maxw + 1))
        qs = sorted(rng.choice(nq, w, replace=False).tolist())
        ops = "".join(rng.choice(list("XYZ"), w))
        recs.append((ops, tuple(qs), float(rng.uniform(-math.pi, math.pi))))
    return wl._pack(recs)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("kw", [
    {"topk": 30},
    {"topk": 12, "mincoeff": 1e-4},
    {"topk": 40, "batched": False},
    {"topk": 25, "max_xweight": 3, "xtrunc_period": 2},
])
def test_topk_random_circuits(ctx, seed, kw):
    rng = np.random.default_rng(500 + seed)
    nq = 10
    g = random_rotations(rng, nq, 50)
    x, z, c = wl.single_pauli("Z", [int(rng.integers(0, nq))], nq)
    ref = pyffo.pauli_propagate(g, 1, x, z, c, **kw)
    dev = ctx.pauli_propagate(g, 1, x, z, c, **kw)
    assert len(dev.c) <= kw["topk"]
    assert_topk_match(dev, ref, f"seed {seed} {kw}")


def test_topk_wide_keys_and_windows(ctx):
    """Commuting layers on 150 qubits (3 words): windows end at top-k points."""
    rng = np.random.default_rng(77)
    nq = 150
    recs = []
    for layer in range(6):
        ops = "Z" if layer % 2 == 0 else "X"
        for _ in range(10):
            w = int(rng.integers(1, 3))
            qs = sorted(rng.choice(nq, w, replace=False).tolist())
            recs.append((ops * w, tuple(qs), float(rng.uniform(-math.pi, math.pi))))
    g = wl._pack(recs)
    W = _abi.words_for_bits(nq)
    x, z, c = wl.single_pauli("Y", [int(rng.integers(0, nq))], nq)
    for topk in (8, 64):
        ref = pyffo.pauli_propagate(g, W, x, z, c, topk=topk)
        dev = ctx.pauli_propagate(g, W, x, z, c, topk=topk)
        assert_topk_match(dev, ref, f"topk {topk}")


def test_topk_with_forced_ties(ctx):
    """One commuting batch of equal-angle X rotations: many terms share each
    magnitude; a single truncation point keeps exactly topk of them."""
    nq = 8
    g = wl._pack([("X", (q,), 0.4) for q in range(nq)])
    x, z, c = wl.single_pauli("ZZZZZZZZ", list(range(nq)), nq)
    for topk in (1, 5, 37, 100, 255, 256, 300):
        ref = pyffo.pauli_propagate(g, 1, x, z, c, topk=topk)
        dev = ctx.pauli_propagate(g, 1, x, z, c, topk=topk)
        assert len(dev.c) == min(topk, 256)
        assert np.array_equal(np.sort(np.abs(dev.c)), np.sort(np.abs(ref.c)))


def test_topk_ties_deterministic(ctx, monkeypatch):
    """Surplus ties at the boundary are ranked by a hash of the key: the kept set is
    the same on every run and for every launch grid (the store's slot order is not:
    blocks reserve their appends in whatever order they finish)."""
    nq = 8
    g = wl._pack([("X", (q,), 0.4) for q in range(nq)] + [("Y", (q,), 0.3) for q in range(nq)])
    x, z, c = wl.single_pauli("ZZZZZZZZ", list(range(nq)), nq)
    runs = []
    for grid in (None, "7", "64", None, "1000"):
        if grid is None:
            monkeypatch.delenv("FFX_WIN_GRID", raising=False)
        else:
            monkeypatch.setenv("FFX_WIN_GRID", grid)
        for topk in (5, 37, 100):
            dev = ctx.pauli_propagate(g, 1, x, z, c, topk=topk, batched=False)
            assert len(dev.c) == topk
            k, v = canon(dev.x, dev.z, dev.c)
            runs.append((topk, grid, k, v))
    for topk in (5, 37, 100):
        same = [(k, v) for (t, _, k, v) in runs if t == topk]
        for k, v in same[1:]:
            assert np.array_equal(k, same[0][0]) and np.array_equal(v, same[0][1]), topk


def test_topk_precisions_and_api(ctx):
    import fastfermion_amd as ff
    rng = np.random.default_rng(9)
    g = random_rotations(rng, 8, 40)
    x, z, c = wl.single_pauli("Z", [3], 8)
    ref = pyffo.pauli_propagate(g, 1, x, z, c, topk=20)
    for prec in (_abi.FP32, _abi.FP32C):
        dev = ctx.pauli_propagate(g, 1, x, z, c, topk=20, precision=prec)
        assert len(dev.c) <= 20
    circ = []
    for r in g:
        ops, qs = [], []
        for q in range(8):
            xb, zb = (int(r["x"][0]) >> q) & 1, (int(r["z"][0]) >> q) & 1
            if xb or zb:
                qs.append(q)
                ops.append({(1, 0): "X", (1, 1): "Y", (0, 1): "Z"}[(xb, zb)])
        circ.append(ff.ROT("".join(ops), qs, float(r["theta"])))
    res = ff.propagate(circ, ff.PauliString("Z3"), topk=20)
    assert len(res) == len(ref.c)
    assert abs(res.overlapwithzero() - pyffo.pauli_expect(ref, [0])) < 1e-12
