"""Layers of single-qubit rotations in one launch (k_layer_async) against one
launch per window / rotation and against the oracle.

A layer replays, orbit by orbit, the same per-rotation operations in the same
order, so FP64 coefficients must be bit-identical to the window path and to
the oracle, and the work counters (term-gate updates, live terms leaving each
rotation, peak) identical too: truncation points inside a layer, X/Y-weight
caps, erased-and-revived slots, complex coefficients, batched and per-gate
schedules, 1..4 key words, X/Y/Z axes mixed in one layer.
"""
from __future__ import annotations

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


def layered_circuit(rng, nq, layers, p_two=0.5):
    """Alternating blocks: two-qubit rotations, then a layer of single-qubit rotations
    with random X/Y/Z axes on a random subset of distinct qubits."""
    spec = []
    for _ in range(layers):
        for _ in range(int(rng.integers(2, 6))):
            a, b = sorted(rng.choice(nq, 2, replace=False).tolist())
            spec.append(("".join(rng.choice(list("XYZ"), 2)), (a, b), float(rng.uniform(-1.2, 1.2))))
        qs = rng.choice(nq, int(rng.integers(5, min(nq, 40) + 1)), replace=False).tolist()
        for q in qs:
            spec.append((str(rng.choice(list("XYZ"))), (q,), float(rng.uniform(-1.2, 1.2))))
    return wl._pack(spec)


def observable(rng, nq, W, terms, complex_coeff=False):
    seen = set()
    xs, zs, cs = [], [], []
    while len(xs) < terms:
        qs = tuple(sorted(rng.choice(nq, int(rng.integers(1, 4)), replace=False).tolist()))
        ops = "".join(rng.choice(list("XYZ"), len(qs)))
        if (ops, qs) in seen:
            continue
        seen.add((ops, qs))
        x, z = _abi.pauli_word(ops, qs, W)
        xs.append(x)
        zs.append(z)
        c = rng.normal()
        cs.append(c + (1j * rng.normal() if complex_coeff else 0))
    return np.array(xs).reshape(-1, W), np.array(zs).reshape(-1, W), np.array(cs, np.complex128)


def both_paths(ctx, monkeypatch, gates, W, x, z, c, **kw):
    monkeypatch.setenv("FFX_LAYERS", "0")
    ref = ctx.pauli_propagate(gates, W, x, z, c, **kw)
    pr = ctx.profile()
    monkeypatch.delenv("FFX_LAYERS")
    got = ctx.pauli_propagate(gates, W, x, z, c, **kw)
    pg = ctx.profile()
    return ref, got, pr, pg


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("batched", [False, True])
def test_layers_bit_exact(ctx, monkeypatch, seed, batched):
    rng = np.random.default_rng(1000 + seed)
    nq = [20, 70, 150, 250][seed]
    W = _abi.words_for_bits(nq)
    gates = layered_circuit(rng, nq, 6)
    x, z, c = observable(rng, nq, W, 3, complex_coeff=seed % 2 == 1)
    kw = dict(maxdegree=6, mincoeff=1e-4, batched=batched)
    ref, got, pr, pg = both_paths(ctx, monkeypatch, gates, W, x, z, c, **kw)
    assert_same_terms(got, ref, exact=True, what=f"layers seed {seed}")
    for k in ("updates", "outs", "peak_terms", "n_gates"):
        assert pg[k] == pr[k], (k, pg[k], pr[k])
    assert pg["run_launches"] < pr["run_launches"]  # the layers did run
    orc = pyffo.pauli_propagate(gates, W, x, z, c, **kw)
    assert_same_terms(got, orc, exact=True, what="layers vs oracle")


@pytest.mark.parametrize("prec", [_abi.FP32, _abi.FP32C])
def test_layers_reduced_precision_identical(ctx, monkeypatch, prec):
    rng = np.random.default_rng(7)
    gates = layered_circuit(rng, 40, 5)
    x, z, c = observable(rng, 40, 1, 2)
    kw = dict(maxdegree=6, mincoeff=1e-4, batched=False, precision=prec)
    ref, got, _, _ = both_paths(ctx, monkeypatch, gates, 1, x, z, c, **kw)
    assert_same_terms(got, ref, exact=True, what="layers reduced precision")


def test_layers_with_xweight_cap(ctx, monkeypatch):
    rng = np.random.default_rng(9)
    gates = layered_circuit(rng, 30, 6)
    x, z, c = observable(rng, 30, 1, 3)
    kw = dict(maxdegree=7, mincoeff=1e-5, batched=False, max_xweight=2, xtrunc_period=3)
    ref, got, pr, pg = both_paths(ctx, monkeypatch, gates, 1, x, z, c, **kw)
    assert_same_terms(got, ref, exact=True, what="layers with X/Y cap")
    assert pg["updates"] == pr["updates"]


def test_kicked_ising_layers(ctx, monkeypatch):
    gates, _ = wl.kicked_ising(6, 5)
    x, z, c = wl.single_pauli("Z", [21], 36)
    for batched in (False, True):
        kw = dict(maxdegree=8, mincoeff=1e-5, batched=batched)
        ref, got, pr, pg = both_paths(ctx, monkeypatch, gates, 1, x, z, c, **kw)
        assert_same_terms(got, ref, exact=True, what=f"kicked Ising batched={batched}")
        assert pg["updates"] == pr["updates"] and pg["peak_terms"] == pr["peak_terms"]


@pytest.mark.parametrize("batched", [False, True])
def test_layers_after_clifford_rotation_runs(ctx, monkeypatch, batched):
    """Pruned mode: runs of Clifford rotations remap erased slots too (keys stay
    distinct), so erased slots can exceed the layer's degree bound; layers copy
    them through.  Layers on and off must agree bit for bit."""
    from fastfermion_amd._native import prune_opts

    gates, _ = wl.kicked_ising(5, 6)
    x, z, c = wl.single_pauli("Z", [12], 25)
    row, col = wl.lattice_geometry(5)
    po = prune_opts(row, col, 12, clifford_tol=1e-12, rot_per_layer=len(gates) // 6)
    kw = dict(maxdegree=8, mincoeff=1e-5, batched=batched)
    monkeypatch.setenv("FFX_LAYERS", "0")
    ref = ctx.pauli_propagate_pruned(gates, 1, x, z, c, po, **kw)
    pr = ctx.profile()
    monkeypatch.delenv("FFX_LAYERS")
    got = ctx.pauli_propagate_pruned(gates, 1, x, z, c, po, **kw)
    pg = ctx.profile()
    assert_same_terms(got, ref, exact=True, what=f"clifford runs + layers batched={batched}")
    assert pg["updates"] == pr["updates"] and pg["run_launches"] < pr["run_launches"]
