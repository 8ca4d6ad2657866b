"""ffx_pauli_propagate_noalloc (the entry point that allocates no device memory)
against ffx_pauli_propagate on the same inputs: identical keys and bit-identical
FP64 coefficients, on two calls in a row (the second with a smaller input, so it
reuses the scratch array the first one filled), plus its capacity errors."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import assert_same_terms
from fastfermion_amd import _abi, workloads as wl
from fastfermion_amd._native import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def check_same(ctx, gates, W, x, z, c, state=None, **kw):
    ref, e_ref = ctx.pauli_propagate(gates, W, x, z, c, state=state or [0] * W, **kw)
    dev, e_dev, n = ctx.pauli_propagate_noalloc(gates, W, x, z, c, state=state or [0] * W, **kw)
    assert n == len(ref) == len(dev)
    assert_same_terms(dev, ref, exact=True, what=str(kw))
    # the device sums <b|O|b> in slot order, and slot order follows the order of the
    # appends' atomics, so two runs of either entry point can differ by rounding
    assert abs(e_dev - e_ref) <= 1e-15 * max(1.0, sum(abs(ref.c)))
    return n


def test_two_calls_second_smaller(ctx):
    # 10x10 kicked Ising, 20 layers, per-gate truncation (windows and layers, 128-bit keys)
    g, _ = wl.kicked_ising(10, 20)
    x, z, c = wl.single_pauli("Z", [55], 100)
    print("noalloc: kicked Ising 10x10x20 per-gate", flush=True)
    n1 = check_same(ctx, g, 2, x, z, c, maxdegree=8, mincoeff=1e-5, batched=False)
    # smaller: the 20-qubit TFIM chain (64-bit keys, exact zeros kept)
    g2 = wl.tfim_chain(20, 5)
    x2, z2, c2 = wl.single_pauli("Z", [10], 20)
    print("noalloc: TFIM n=20", flush=True)
    n2 = check_same(ctx, g2, 1, x2, z2, c2, maxdegree=6)
    assert n2 < n1 or len(g2) < len(g)


def test_precisions_complex_and_clifford(ctx):
    rng = np.random.default_rng(5)
    g = np.concatenate([wl._pack([("XY", (0, 1), 0.3), ("Z", (2,), 0.7)]), _abi.clifford_record(_abi.H, 1),
                        wl._pack([("YZ", (1, 2), -0.4)])] * 6)
    x, z, c = wl.single_pauli("ZX", [0, 2], 3, coeff=0.5 - 0.25j)
    print("noalloc: Clifford + complex", flush=True)
    check_same(ctx, g, 1, x, z, c, mincoeff=1e-6)
    # 6x6 kicked Ising, 3 layers, reference schedule: peak 44215 live terms
    g2, _ = wl.kicked_ising(6, 3)
    x2, z2, c2 = wl.single_pauli("Z", [18], 36)
    for prec in (_abi.FP64, _abi.FP32, _abi.FP32C):
        for batched in (True, False):
            print(f"noalloc: kicked Ising 6x6x3 precision {prec} batched {batched}", flush=True)
            check_same(ctx, g2, 1, x2, z2, c2, maxdegree=8, mincoeff=1e-5, batched=batched, precision=prec)


def test_expectation_only_and_small_output(ctx):
    g, _ = wl.kicked_ising(6, 4)
    x, z, c = wl.single_pauli("Z", [14], 36)
    ref, e_ref = ctx.pauli_propagate(g, 1, x, z, c, maxdegree=8, mincoeff=1e-5, state=[0])
    res, e, n = ctx.pauli_propagate_noalloc(g, 1, x, z, c, maxdegree=8, mincoeff=1e-5, state=[0], out_cap=0)
    assert res is None and n == len(ref) and abs(e - e_ref) <= 1e-15
    with pytest.raises(RuntimeError, match="output arrays hold"):
        ctx.pauli_propagate_noalloc(g, 1, x, z, c, maxdegree=8, mincoeff=1e-5, out_cap=len(ref) - 1)
    assert ctx.last_noalloc_terms == len(ref)


def test_rejects_stores_beyond_the_scratch(ctx):
    # the reference schedule of the 10x10 circuit peaks at 1.36e6 live terms: more
    # than the 524288 slots the scratch array holds for 128-bit FP64 keys
    g, _ = wl.kicked_ising(10, 20)
    x, z, c = wl.single_pauli("Z", [55], 100)
    print("noalloc: over capacity", flush=True)
    with pytest.raises(RuntimeError, match="slots its scratch holds"):
        ctx.pauli_propagate_noalloc(g, 2, x, z, c, maxdegree=8, mincoeff=1e-5, batched=True, out_cap=0)
    # an observable larger than the scratch array is rejected before any launch
    n = 1 << 21
    xs = np.zeros((n, 1), np.uint64)
    zs = np.arange(1, n + 1, dtype=np.uint64).reshape(n, 1)
    with pytest.raises(RuntimeError, match="exceed its scratch"):
        ctx.pauli_propagate_noalloc(wl.tfim_chain(20, 1), 1, xs, zs, np.ones(n), out_cap=0)
    # the context still works afterwards, with and without the scratch array
    g2 = wl.tfim_chain(20, 5)
    x2, z2, c2 = wl.single_pauli("Z", [10], 20)
    check_same(ctx, g2, 1, x2, z2, c2, maxdegree=6)
