"""Layers of rotations with pairwise disjoint supports in one launch
(k_dlayer_async) against one launch per window / rotation and against the
oracle.

Runs of at least five consecutive rotations whose axes share no qubit (Pauli)
or no Majorana index (Majorana strings of even degree) become one launch that
replays every orbit's rotations in layer order with the per-rotation
operations, so FP64 coefficients must be bit-identical to the oracle's and the
term-gate update count identical: Fermi-Hubbard Trotter circuits in both term
orders (bond order: rows of vertical bonds form layers; layered order: every
group is one), 1..3 key words, truncation after every Majorana gate, partner
degree limits from -1 to 8, real and complex observables; Pauli circuits of
random multi-qubit axes on disjoint supports mixed with overlapping rotations,
batched and per-gate truncation schedules, X/Y-weight caps, 1..4 key words.
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


def _launches(ctx, monkeypatch, run):
    """(result, profile) with disjoint layers off, then on."""
    monkeypatch.setenv("FFX_NO_DLAYERS", "1")
    ref = run()
    pr = ctx.profile()
    monkeypatch.delenv("FFX_NO_DLAYERS")
    got = run()
    pg = ctx.profile()
    return ref, got, pr, pg


@pytest.mark.parametrize("lx,ly,order,steps", [(2, 2, "layers", 6), (4, 4, "layers", 4), (4, 4, "bonds", 4),
                                               (5, 5, "layers", 3), (6, 6, "layers", 2), (6, 6, "bonds", 2)])
def test_fermi_hubbard_disjoint_layers_bit_exact(ctx, monkeypatch, lx, ly, order, steps):
    g, (m, c) = wl.fermi_hubbard(lx, ly, steps=steps, order=order)
    W = m.shape[1]
    ref = pyffo.majorana_propagate(g, W, m, c, maxdegree=8, mincoeff=1e-4)
    off, dev, pr, pg = _launches(ctx, monkeypatch,
                                 lambda: ctx.majorana_propagate(g, W, m, c, maxdegree=8, mincoeff=1e-4))
    print(f"FH {lx}x{ly} {order} x{steps}: {len(dev)} terms, launches {pr['run_launches']} -> "
          f"{pg['run_launches']}", flush=True)
    assert_same_terms(dev, ref, exact=True, what=f"FH {lx}x{ly} {order}")
    assert_same_terms(off, ref, exact=True, what=f"FH {lx}x{ly} {order} without layers")
    assert pg["updates"] == ref.updates == pr["updates"]
    assert pg["run_launches"] < pr["run_launches"]
    occ = sum(1 << (2 * s) for s in range(lx * ly) if s % 2 == 0)
    occw = [(occ >> (64 * w)) & ((1 << 64) - 1) for w in range(W)]
    assert abs(ctx.expect_last(occw) - pyffo.majorana_expect(ref, occw)) < 1e-15


@pytest.mark.parametrize("maxdegree", [-1, 0, 2, 4, 6])
def test_fermi_hubbard_degree_limits(ctx, maxdegree):
    g, (m, c) = wl.fermi_hubbard(4, 4, steps=3, order="layers")
    ref = pyffo.majorana_propagate(g, 1, m, c, maxdegree=maxdegree, mincoeff=1e-5)
    dev = ctx.majorana_propagate(g, 1, m, c, maxdegree=maxdegree, mincoeff=1e-5)
    assert_same_terms(dev, ref, exact=True, what=f"FH maxdegree {maxdegree}")
    assert ctx.profile()["updates"] == ref.updates


@pytest.mark.parametrize("real", [True, False])
def test_fermi_hubbard_real_and_complex_observables(ctx, real):
    """A real observable runs the one-plane kernels, a complex one the two-plane ones."""
    g, _ = wl.fermi_hubbard(4, 4, steps=3, order="layers")
    rng = np.random.default_rng(5)
    m = np.zeros((6, 1), np.uint64)
    for i in range(6):
        for a in rng.choice(64, 2 * int(rng.integers(1, 3)), replace=False):
            m[i, 0] |= np.uint64(1) << np.uint64(int(a))
    c = rng.normal(size=6) + (0 if real else 1j * rng.normal(size=6))
    ref = pyffo.majorana_propagate(g, 1, m, c, maxdegree=8, mincoeff=1e-5)
    dev = ctx.majorana_propagate(g, 1, m, c, maxdegree=8, mincoeff=1e-5)
    assert_same_terms(dev, ref, exact=True, what=f"FH observable real={real}")


def majorana_mixed_circuit(rng, n, blocks):
    """Runs of even-degree strings on disjoint index sets (lengths 3..40), with odd-degree
    and overlapping strings between them."""
    spec = []
    for _ in range(blocks):
        free = rng.permutation(n).tolist()
        for _ in range(int(rng.integers(3, 41))):
            d = int(rng.choice([2, 2, 4, 6]))
            if len(free) < d:
                break
            idx = sorted(free[:d])
            free = free[d:]
            spec.append((idx, float(rng.uniform(-1.0, 1.0)), (d >> 1) & 1))
        for _ in range(int(rng.integers(1, 4))):
            d = int(rng.integers(1, 5))
            idx = sorted(rng.choice(n, d, replace=False).tolist())
            spec.append((idx, float(rng.uniform(-1.0, 1.0)), (d >> 1) & 1))
    return wl.pack_majorana(spec)


@pytest.mark.parametrize("n,seed", [(40, 0), (128, 1), (190, 2), (250, 3)])
def test_majorana_random_disjoint_runs(ctx, monkeypatch, n, seed):
    rng = np.random.default_rng(seed)
    W = _abi.words_for_bits(n)
    g = majorana_mixed_circuit(rng, n, 8)
    m = np.zeros((5, W), np.uint64)
    for i in range(5):
        for a in rng.choice(n, 2 * int(rng.integers(1, 3)), replace=False):
            m[i, int(a) >> 6] |= np.uint64(1) << np.uint64(int(a) & 63)
    c = rng.normal(size=5) + 1j * rng.normal(size=5)
    ref = pyffo.majorana_propagate(g, W, m, c, maxdegree=8, mincoeff=1e-4)
    off, dev, pr, pg = _launches(ctx, monkeypatch,
                                 lambda: ctx.majorana_propagate(g, W, m, c, maxdegree=8, mincoeff=1e-4))
    assert_same_terms(dev, ref, exact=True, what=f"Majorana n={n} seed={seed}")
    assert pg["updates"] == ref.updates
    assert pg["run_launches"] < pr["run_launches"]


def pauli_disjoint_circuit(rng, nq, blocks):
    """Runs of random 1..4-qubit axes on disjoint qubit sets (lengths 3..30), with a few
    overlapping two-qubit rotations between them."""
    spec = []
    for _ in range(blocks):
        free = rng.permutation(nq).tolist()
        for _ in range(int(rng.integers(3, 31))):
            d = int(rng.integers(1, 5))
            if len(free) < d:
                break
            qs = tuple(sorted(free[:d]))
            free = free[d:]
            spec.append(("".join(rng.choice(list("XYZ"), d)), qs, float(rng.uniform(-1.2, 1.2))))
        for _ in range(int(rng.integers(1, 4))):
            a, b = sorted(rng.choice(nq, 2, replace=False).tolist())
            spec.append(("".join(rng.choice(list("XYZ"), 2)), (a, b), float(rng.uniform(-1.2, 1.2))))
    return wl._pack(spec)


def pauli_observable(rng, nq, W, terms):
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
        cs.append(rng.normal() + 1j * rng.normal())
    return np.array(xs).reshape(-1, W), np.array(zs).reshape(-1, W), np.array(cs, np.complex128)


@pytest.mark.parametrize("nq,seed", [(40, 0), (100, 1), (150, 2), (250, 3)])
@pytest.mark.parametrize("batched", [True, False])
def test_pauli_disjoint_runs_bit_exact(ctx, monkeypatch, nq, seed, batched):
    rng = np.random.default_rng(10 + seed)
    W = _abi.words_for_bits(nq)
    g = pauli_disjoint_circuit(rng, nq, 12)
    x, z, c = pauli_observable(rng, nq, W, 20)
    kw = dict(maxdegree=8, mincoeff=1e-6, batched=batched)
    ref = pyffo.pauli_propagate(g, W, x, z, c, **kw)
    off, dev, pr, pg = _launches(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, W, x, z, c, **kw))
    print(f"Pauli nq={nq} batched={batched}: {len(dev)} terms, launches {pr['run_launches']} -> "
          f"{pg['run_launches']}", flush=True)
    assert_same_terms(dev, ref, exact=True, what=f"Pauli disjoint nq={nq} batched={batched}")
    assert_same_terms(off, ref, exact=True, what="without disjoint layers")
    assert pg["updates"] == ref.updates
    assert pg["run_launches"] < pr["run_launches"]


def test_pauli_disjoint_runs_with_xweight_cap(ctx, monkeypatch):
    rng = np.random.default_rng(21)
    nq, W = 100, 2
    g = pauli_disjoint_circuit(rng, nq, 10)
    x, z, c = pauli_observable(rng, nq, W, 6)
    kw = dict(maxdegree=8, mincoeff=1e-5, batched=True, max_xweight=3)
    ref = pyffo.pauli_propagate(g, W, x, z, c, **kw)
    _, dev, pr, pg = _launches(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, W, x, z, c, **kw))
    assert_same_terms(dev, ref, exact=True, what="Pauli disjoint with X/Y cap")
    assert pg["updates"] == ref.updates


@pytest.mark.parametrize("prec", [_abi.FP32, _abi.FP32C])
def test_disjoint_layers_reduced_precision_identical(ctx, monkeypatch, prec):
    """In FP32 and FP32 compensated the layer launch performs the same operations in
    the same order as the window launches: results identical with layers on and off."""
    g, (m, c) = wl.fermi_hubbard(4, 4, steps=4, order="layers")
    off, dev, pr, pg = _launches(ctx, monkeypatch,
                                 lambda: ctx.majorana_propagate(g, 1, m, c, maxdegree=8, mincoeff=1e-4,
                                                                precision=prec))
    assert_same_terms(dev, off, exact=True, what=f"FH prec={prec}")


def pauli_commuting_circuit(rng, nq, blocks):
    """Runs of 3..12 commuting rotations with overlapping supports (Z-type or X-type
    strings of weight 2..3 on nearby qubits, some closing cycles, i.e. linearly
    dependent on earlier axes of the run), with single-qubit and random two-qubit
    rotations between the runs."""
    spec = []
    for _ in range(blocks):
        op = str(rng.choice(["Z", "X"]))
        base = int(rng.integers(0, nq - 12))
        for _ in range(int(rng.integers(3, 13))):
            w = int(rng.integers(2, 4))
            qs = tuple(sorted((base + rng.choice(12, w, replace=False)).tolist()))
            spec.append((op * w, qs, float(rng.uniform(-1.2, 1.2))))
        for _ in range(int(rng.integers(1, 4))):
            if rng.random() < 0.5:
                spec.append((str(rng.choice(list("XYZ"))), (int(rng.integers(nq)),), float(rng.uniform(-1.2, 1.2))))
            else:
                a, b = sorted(rng.choice(nq, 2, replace=False).tolist())
                spec.append(("".join(rng.choice(list("XYZ"), 2)), (a, b), float(rng.uniform(-1.2, 1.2))))
    return wl._pack(spec)


@pytest.mark.parametrize("nq,seed,blocks", [(40, 0, 12), (100, 1, 14), (200, 2, 14)])
@pytest.mark.parametrize("batched", [True, False])
def test_pauli_general_layers_bit_exact(ctx, monkeypatch, nq, seed, blocks, batched):
    """General layers (up to eight commuting rotations on overlapping supports, one
    launch) against the oracle, and against the run without them."""
    rng = np.random.default_rng(40 + seed)
    W = _abi.words_for_bits(nq)
    g = pauli_commuting_circuit(rng, nq, blocks)
    x, z, c = pauli_observable(rng, nq, W, 20)
    kw = dict(maxdegree=8, mincoeff=1e-6, batched=batched)
    ref = pyffo.pauli_propagate(g, W, x, z, c, **kw)
    off, dev, pr, pg = _launches(ctx, monkeypatch, lambda: ctx.pauli_propagate(g, W, x, z, c, **kw))
    print(f"Pauli general nq={nq} batched={batched}: {len(dev)} terms, launches {pr['run_launches']} -> "
          f"{pg['run_launches']}", flush=True)
    assert_same_terms(dev, ref, exact=True, what=f"Pauli general nq={nq} batched={batched}")
    assert_same_terms(off, ref, exact=True, what="without general layers")
    assert pg["updates"] == ref.updates
    assert pg["run_launches"] < pr["run_launches"]


@pytest.mark.parametrize("seed", range(3))
def test_majorana_general_layers_bit_exact(ctx, monkeypatch, seed):
    """Degree-4 strings made of two index pairs {2i, 2i+1}: any two overlap in 0 or 2
    indices, so runs of them commute with overlapping supports (three strings on
    the pairs of a triangle are linearly dependent and end a run)."""
    rng = np.random.default_rng(60 + seed)
    n, W = 120, 2
    spec = []
    for _ in range(12):
        base = int(rng.integers(0, n // 2 - 6))
        for _ in range(int(rng.integers(3, 12))):
            p1, p2 = sorted((base + rng.choice(6, 2, replace=False)).tolist())
            spec.append(([2 * p1, 2 * p1 + 1, 2 * p2, 2 * p2 + 1], float(rng.uniform(-1.0, 1.0)), 0))
        for _ in range(int(rng.integers(1, 3))):
            d = int(rng.integers(1, 5))
            idx = sorted(rng.choice(n, d, replace=False).tolist())
            spec.append((idx, float(rng.uniform(-1.0, 1.0)), (d >> 1) & 1))
    g = wl.pack_majorana(spec)
    m = np.zeros((6, W), np.uint64)
    for i in range(6):
        for a in rng.choice(n, 2 * int(rng.integers(1, 3)), replace=False):
            m[i, int(a) >> 6] |= np.uint64(1) << np.uint64(int(a) & 63)
    c = rng.normal(size=6) + 1j * rng.normal(size=6)
    ref = pyffo.majorana_propagate(g, W, m, c, maxdegree=8, mincoeff=1e-5)
    off, dev, pr, pg = _launches(ctx, monkeypatch,
                                 lambda: ctx.majorana_propagate(g, W, m, c, maxdegree=8, mincoeff=1e-5))
    assert_same_terms(dev, ref, exact=True, what=f"Majorana general seed={seed}")
    assert_same_terms(off, ref, exact=True, what="without general layers")
    assert pg["updates"] == ref.updates
    assert pg["run_launches"] < pr["run_launches"]
