"""Device engine (libffamd.so through its C ABI) against the CPU oracle.

FP64 results must have the oracle's key set and bit-identical coefficients: the
device computes every per-gate value with the same operations and rounding as
the reference (c*cos, +-(c*sin), one add), and erased/revived slots follow the
reference's erase/re-insert semantics.  FP32 variants are checked against the
same oracle within stated tolerances.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import pyffo
from _helpers import assert_close_terms, assert_same_terms, rel_err
from fastfermion_amd import _abi, workloads as wl
from fastfermion_amd._native import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def rot(ops, qubits, theta):
    return wl._pack([(ops, tuple(qubits), theta)])


def cat(*parts):
    return np.concatenate([p for p in parts if len(p)]) if any(len(p) for p in parts) else \
        np.zeros(0, _abi.PAULI_GATE_DTYPE)


def tfim(n, dt):
    g = [("ZZ", (i, (i + 1) % n), -2 * dt) for i in range(n)] + [("X", (i,), -2 * dt) for i in range(n)]
    return wl._pack(g)


def heisenberg(n, dt):
    g = []
    for i in range(n):
        j = (i + 1) % n
        g += [("XX", (i, j), 2 * dt), ("YY", (i, j), 2 * dt), ("ZZ", (i, j), 2 * dt)]
    return wl._pack(g)


def obs(spec, nq):
    """spec: list of (ops, qubits, coeff) -> term-major keys."""
    W = _abi.words_for_bits(nq)
    xs, zs, cs = [], [], []
    for ops, qs, c in spec:
        x, z = _abi.pauli_word(ops, qs, W)
        xs.append(x)
        zs.append(z)
        cs.append(c)
    return np.array(xs, np.uint64).reshape(-1, W), np.array(zs, np.uint64).reshape(-1, W), np.array(cs, np.complex128)


def both(ctx, gates, W, x, z, c, **kw):
    ref = pyffo.pauli_propagate(gates, W, x, z, c, **kw)
    prec = kw.pop("precision", _abi.FP64)
    dev = ctx.pauli_propagate(gates, W, x, z, c, precision=prec, **kw)
    return ref, dev


CASES = [
    ("tfim6", lambda: tfim(6, 0.1), 5, [("Z", [0], 1.0)], 6, {}),
    ("tfim6_thresh", lambda: tfim(6, 0.1), 5, [("Z", [0], 1.0)], 6, {"mincoeff": 1e-3}),
    ("tfim6_maxdeg", lambda: tfim(6, 0.1), 5, [("Z", [0], 1.0)], 6, {"maxdegree": 3}),
    # maxdegree is compared literally (src/pauli/propagate.h:82): no partner survives -1 or 0
    ("tfim6_maxdeg_neg", lambda: tfim(6, 0.1), 5, [("Z", [0], 1.0), ("ZZ", [1, 2], 0.5)], 6, {"maxdegree": -1}),
    ("tfim6_maxdeg0", lambda: tfim(6, 0.1), 5, [("Z", [0], 1.0)], 6, {"maxdegree": 0}),
    ("heis4", lambda: heisenberg(4, 0.03), 5, [("Z", [0], 1.0)], 4, {}),
    ("poly_obs", lambda: tfim(4, 0.05), 10, [("Z", [0], 1.0), ("Z", [1], 1.0)], 4, {}),
    ("identity_rot", lambda: rot("X", [0], 0.0), 10, [("Z", [0], 1.0)], 1, {}),
    ("all_commuting", lambda: rot("ZZ", [0, 1], 0.3), 10, [("Z", [0], 1.0)], 2, {}),
    ("xweight", lambda: tfim(6, 0.1), 5, [("Z", [0], 1.0)], 6, {"max_xweight": 2, "xtrunc_period": 3}),
    ("complex_coeffs", lambda: heisenberg(5, 0.07), 4, [("X", [0], 0.5 + 0.25j), ("YZ", [1, 2], -1j)], 5,
     {"mincoeff": 1e-4}),
]


@pytest.mark.parametrize("batched", [True, False])
@pytest.mark.parametrize("name,gf,steps,spec,nq,kw", CASES, ids=[c[0] for c in CASES])
def test_reference_suite_cases_bit_exact(ctx, name, gf, steps, spec, nq, kw, batched):
    """The circuits of the reference's omp/sharded/batched suites
    (test/omp_test.py:23-108, test/batched_test.py:26-83)."""
    g = np.concatenate([gf()] * steps)
    x, z, c = obs(spec, nq)
    ref, dev = both(ctx, g, 1, x, z, c, batched=batched, **kw)
    assert_same_terms(dev, ref, exact=True, what=name)


def test_empty_circuit(ctx):
    x, z, c = obs([("Z", [0], 1.0)], 1)
    ref, dev = both(ctx, np.zeros(0, _abi.PAULI_GATE_DTYPE), 1, x, z, c)
    assert_same_terms(dev, ref, exact=True)


def test_empty_observable(ctx):
    x = np.zeros((0, 1), np.uint64)
    dev = ctx.pauli_propagate(tfim(4, 0.1), 1, x, x, np.zeros(0))
    assert len(dev) == 0


def test_mixed_clifford_rotations(ctx):
    c1 = np.concatenate([_abi.clifford_record(_abi.H, 0), rot("Z", [0], 0.5),
                         _abi.clifford_record(_abi.CNOT, 0, 1), rot("XX", [0, 1], 0.3)])
    x, z, c = obs([("Z", [0], 1.0)], 2)
    for batched in (True, False):
        ref, dev = both(ctx, np.concatenate([c1] * 5), 1, x, z, c, batched=batched)
        assert_same_terms(dev, ref, exact=True, what=f"batched={batched}")


def test_clifford_only_all_kinds(ctx):
    g = np.concatenate([_abi.clifford_record(k, 0, 1) for k in (_abi.H, _abi.S, _abi.CNOT, _abi.SWAP, _abi.CZ)]
                       + [_abi.clifford_record(_abi.CNOT, 2, 0), _abi.clifford_record(_abi.S, 2)])
    rng = np.random.default_rng(3)
    spec = [("".join(rng.choice(list("IXYZ"), 3)), [0, 1, 2], complex(rng.normal(), rng.normal())) for _ in range(12)]
    x, z, c = obs(spec, 3)
    keys = {}
    for i in range(len(c)):
        keys[(int(x[i, 0]), int(z[i, 0]))] = c[i]
    x = np.array([[k[0]] for k in keys], np.uint64)
    z = np.array([[k[1]] for k in keys], np.uint64)
    c = np.array(list(keys.values()))
    ref, dev = both(ctx, g, 1, x, z, c)
    assert_same_terms(dev, ref, exact=True)


def random_circuit(rng, nq, ngates, p_cliff=0.15, maxw=3):
    recs = []
    for _ in range(ngates):
        if rng.random() < p_cliff:
            kind = int(rng.integers(1, 6))
            a, b = rng.choice(nq, 2, replace=False)
            recs.append(_abi.clifford_record(kind, int(a), int(b)))
        else:
            w = int(rng.integers(1, maxw + 1))
            qs = sorted(rng.choice(nq, w, replace=False).tolist())
            ops = "".join(rng.choice(list("XYZ"), w))
            recs.append(rot(ops, qs, float(rng.uniform(-math.pi, math.pi))))
    return np.concatenate(recs)


@pytest.mark.parametrize("seed", range(6))
def test_random_circuits_wide_keys(ctx, seed):
    """Random rotations/Cliffords over 130..250 qubits (3-4 words per mask)."""
    rng = np.random.default_rng(100 + seed)
    nq = int(rng.integers(130, 250))
    W = _abi.words_for_bits(nq)
    g = random_circuit(rng, nq, 60)
    spec = [("Z", [int(q)], 1.0) for q in rng.choice(nq, 3, replace=False)]
    x, z, c = obs(spec, nq)
    for kw in ({"maxdegree": 5, "mincoeff": 1e-6}, {"maxdegree": 4, "batched": False, "mincoeff": 1e-4}):
        ref, dev = both(ctx, g, W, x, z, c, **kw)
        assert_same_terms(dev, ref, exact=True, what=str(kw))


def test_tfim_chain_20(ctx):
    """n=20 open-chain TFIM, 5 steps, maxdegree 6, FP64: bit-exact terms and <Z_10>."""
    g = wl.tfim_chain()
    x, z, c = wl.single_pauli("Z", [10], 20)
    ref = pyffo.pauli_propagate(g, 1, x, z, c, maxdegree=6)
    dev, ex = ctx.pauli_propagate(g, 1, x, z, c, maxdegree=6, state=[0])
    assert_same_terms(dev, ref, exact=True)
    assert rel_err(ex, pyffo.pauli_expect(ref, [0])) < 1e-12


@pytest.mark.parametrize("batched", [False, True])
def test_kicked_ising_6x6(ctx, batched):
    g, _ = wl.kicked_ising(6, 6)
    x, z, c = wl.single_pauli("Z", [21], 36)
    ref = pyffo.pauli_propagate(g, 1, x, z, c, maxdegree=8, mincoeff=1e-5, batched=batched)
    dev, ex = ctx.pauli_propagate(g, 1, x, z, c, maxdegree=8, mincoeff=1e-5, batched=batched, state=[0])
    assert_same_terms(dev, ref, exact=True)
    assert abs(ex - pyffo.pauli_expect(ref, [0])) < 1e-12


def test_product_state_expectation(ctx):
    g = tfim(8, 0.13)
    g = np.concatenate([g] * 4)
    x, z, c = obs([("Z", [0], 1.0), ("ZZ", [2, 3], 0.5), ("X", [5], 0.25)], 8)
    for b in (0, 0b10110101, 0xFF):
        ref = pyffo.pauli_propagate(g, 1, x, z, c)
        _, ex = ctx.pauli_propagate(g, 1, x, z, c, state=[b])
        assert abs(ex - pyffo.pauli_expect(ref, [b])) < 1e-12


@pytest.mark.parametrize("prec,tol", [(_abi.FP32C, 1e-12), (_abi.FP32, 1e-5)])
def test_fp32_variants(ctx, prec, tol):
    """FP32 storage (plain and compensated) vs the FP64 oracle: absolute
    coefficient error <= tol, key sets equal up to keys at the threshold."""
    g, _ = wl.kicked_ising(5, 5, theta_lo=0.3, theta_hi=0.7)
    x, z, c = wl.single_pauli("Z", [12], 25)
    kw = dict(maxdegree=8, mincoeff=1e-7, batched=False)
    ref = pyffo.pauli_propagate(g, 1, x, z, c, **kw)
    dev, ex = ctx.pauli_propagate(g, 1, x, z, c, precision=prec, state=[0], **kw)
    assert_close_terms(dev, ref, tol, mincoeff=1e-7, what=f"precision {prec}")
    assert abs(ex - pyffo.pauli_expect(ref, [0])) < 100 * tol


# ---------------------------------------------------------------------------
# Majorana
# ---------------------------------------------------------------------------

def mrot(idx, theta):
    r = (len(idx) >> 1) & 1
    return wl.pack_majorana([(idx, theta, r)])


def random_majorana(rng, nm, ngates, maxdeg=4):
    gs = []
    for _ in range(ngates):
        d = int(rng.integers(1, maxdeg + 1))
        idx = sorted(rng.choice(nm, d, replace=False).tolist())
        gs.append((idx, float(rng.uniform(-2, 2)), (d >> 1) & 1))
    return wl.pack_majorana(gs)


def mobs(spec, nm):
    W = _abi.words_for_bits(nm)
    ms, cs = [], []
    for idx, c in spec:
        m = np.zeros(W, np.uint64)
        for a in idx:
            m[a >> 6] |= np.uint64(1) << np.uint64(a & 63)
        ms.append(m)
        cs.append(c)
    return np.array(ms, np.uint64).reshape(-1, W), np.array(cs, np.complex128)


@pytest.mark.parametrize("seed", range(5))
def test_majorana_random_bit_exact(ctx, seed):
    rng = np.random.default_rng(7 + seed)
    nm = [12, 40, 100, 150, 250][seed]
    W = _abi.words_for_bits(nm)
    g = random_majorana(rng, nm, 40)
    spec = [([0, 1], 0.5j), ([2], -0.3), ([1, 3, 4, 7], 0.7)] if nm > 8 else [([0, 1], 1j)]
    m, c = mobs(spec, nm)
    for maxdeg, minc in ((-1, 0.0), (4, 1e-5), (6, 1e-3)):
        ref = pyffo.majorana_propagate(g, W, m, c, maxdegree=maxdeg, mincoeff=minc)
        dev = ctx.majorana_propagate(g, W, m, c, maxdegree=maxdeg, mincoeff=minc)
        assert_same_terms(dev, ref, exact=True, what=f"maxdeg={maxdeg} mincoeff={minc}")


def test_majorana_fock_expectation(ctx):
    rng = np.random.default_rng(11)
    g = random_majorana(rng, 16, 30, maxdeg=4)
    m, c = mobs([([0, 1], 0.5j), ([], 0.5), ([2, 3, 4, 5], -0.25)], 16)
    ref = pyffo.majorana_propagate(g, 1, m, c, maxdegree=6, mincoeff=1e-6)
    for occ in (0, 0b1011, 0xFF):
        _, ex = ctx.majorana_propagate(g, 1, m, c, maxdegree=6, mincoeff=1e-6, occ=[occ])
        want = pyffo.majorana_expect(ref, [occ])
        assert abs(ex - want) < 1e-12, (occ, ex, want)


def test_capacity_growth_and_resume():
    """A fresh context starts with a small store; the 6x6 kicked-Ising run
    (peak ~1.4e6 live terms) overflows it several times and must resume each
    time from the overflowing rotation with an identical result."""
    g, _ = wl.kicked_ising(6, 4)
    x, z, c = wl.single_pauli("Z", [21], 36)
    ref = pyffo.pauli_propagate(g, 1, x, z, c, maxdegree=8, mincoeff=1e-5, batched=True)
    fresh = Context(0)
    try:
        dev = fresh.pauli_propagate(g, 1, x, z, c, maxdegree=8, mincoeff=1e-5, batched=True)
        assert_same_terms(dev, ref, exact=True)
        assert fresh.profile()["peak_terms"] <= ref.peak_terms
        assert fresh.profile()["updates"] == ref.updates
    finally:
        fresh.close()


def test_step_engine_matches_async_engine(ctx, monkeypatch):
    """The cross-check engine (host sync after every rotation) and the default
    asynchronous engine produce identical terms and work counters."""
    g, _ = wl.kicked_ising(5, 4)
    x, z, c = wl.single_pauli("Z", [12], 25)
    kw = dict(maxdegree=8, mincoeff=1e-6, batched=False)
    a = ctx.pauli_propagate(g, 1, x, z, c, **kw)
    pa = ctx.profile()
    monkeypatch.setenv("FFX_ENGINE", "steps")
    b = ctx.pauli_propagate(g, 1, x, z, c, **kw)
    pb = ctx.profile()
    assert_same_terms(a, b, exact=True)
    assert pa["updates"] == pb["updates"] and pa["peak_terms"] == pb["peak_terms"]


def test_layout_switch_after_large_result():
    """Regression: a Majorana run leaving many slots, then a Pauli run with a
    different store layout on the same context; the store is re-laid out and
    the index rebuilt from the new slot count only."""
    c2 = Context(0)
    try:
        rng = np.random.default_rng(5)
        gm = random_majorana(rng, 16, 40, maxdeg=4)
        m, cm = mobs([([0, 1], 0.5j), ([2, 3, 4, 5], 0.3)], 16)
        big = c2.majorana_propagate(gm, 1, m, cm, maxdegree=8)
        assert len(big) > 2048
        x, z, c = obs([("Z", [0], 1.0)], 6)
        g = np.concatenate([tfim(6, 0.1)] * 3)
        ref = pyffo.pauli_propagate(g, 1, x, z, c)
        dev = c2.pauli_propagate(g, 1, x, z, c)
        assert_same_terms(dev, ref, exact=True)
    finally:
        c2.close()


def test_compaction_interval_does_not_change_results(ctx, monkeypatch):
    """With truncation after every rotation the device-decided compaction check
    runs after every FFX_COMPACT_EVERY-th truncation point: checking every 4th
    compacts several times, the default interval rarely; terms, coefficients and
    the work counters must not depend on it (compaction keeps slot order)."""
    gates, _ = wl.kicked_ising(6, 5)
    x, z, c = wl.single_pauli("Z", [21], 36)
    kw = dict(maxdegree=8, mincoeff=1e-4, batched=False)
    monkeypatch.setenv("FFX_COMPACT_EVERY", "4")
    often, e_often = ctx.pauli_propagate(gates, 1, x, z, c, want_expect=True, **kw)
    p_often = ctx.profile()
    monkeypatch.delenv("FFX_COMPACT_EVERY")
    dflt, e_dflt = ctx.pauli_propagate(gates, 1, x, z, c, want_expect=True, **kw)
    p_dflt = ctx.profile()
    assert p_often["compactions"] > p_dflt["compactions"]
    assert_same_terms(often, dflt, exact=True, what="compaction interval")
    for k in ("updates", "outs", "peak_terms"):
        assert p_often[k] == p_dflt[k], (k, p_often[k], p_dflt[k])
    assert abs(e_often - e_dflt) < 1e-12
    orc = pyffo.pauli_propagate(gates, 1, x, z, c, **kw)
    assert_same_terms(dflt, orc, exact=True, what="compaction interval vs oracle")


@pytest.mark.parametrize("maxdegree", [-1, 1])
def test_kicked_ising_literal_maxdegree(ctx, maxdegree):
    """Window and layer kernels with a bound below every term's degree: only the
    cos-scaled copies survive, identical to the oracle."""
    g, _ = wl.kicked_ising(L=4, layers=3)
    x, z, c = wl.single_pauli("ZZ", [5, 6], 16)
    for batched in (True, False):
        ref, dev = both(ctx, g, 1, x, z, c, maxdegree=maxdegree, batched=batched)
        assert_same_terms(dev, ref, exact=True, what=f"maxdegree={maxdegree} batched={batched}")
    assert len(ref) == 1  # every partner of Z5 Z6 has degree >= 2


def test_majorana_literal_maxdegree(ctx):
    rng = np.random.default_rng(11)
    g = random_majorana(rng, 12, 30)
    m, c = mobs([([0, 1], 1.0), ([2, 5], 0.5j)], 12)
    for md in (-1, 0, 2):
        ref = pyffo.majorana_propagate(g, 1, m, c, maxdegree=md)
        dev = ctx.majorana_propagate(g, 1, m, c, maxdegree=md)
        assert_same_terms(dev, ref, exact=True, what=f"maxdegree={md}")


@pytest.mark.parametrize("batched", [True, False])
def test_complex_threshold_boundary(ctx, batched):
    """|c| <= mincoeff for a complex coefficient (src/pauli/truncate.h:16-18) at
    the boundary: c = (3 + 4i) 2^-20 has |c| = 5 * 2^-20 exactly.  mincoeff equal
    to it erases the term, one ulp below keeps it, in the one-rotation kernel and
    inside a window of commuting rotations.  (Magnitudes that are not exactly
    representable can round differently in glibc's hypot and the device's; the
    reference's std::abs is glibc's hypot, which is not correctly rounded.)"""
    e = 2.0 ** -20
    x, z, c = obs([("Z", [0], (3 + 4j) * e), ("Z", [1], 0.5 + 0.5j), ("X", [2], 1.0)], 4)
    # a truncation point that leaves the boundary term untouched: Z rotations commute with Z0
    single = rot("Z", [3], 0.3)
    window = np.concatenate([rot("Z", [3], 0.3), rot("ZZ", [1, 3], 0.2), rot("Z", [2], 0.4), rot("ZZ", [0, 1], 0.1)])
    for g in (single, window):
        for mc, kept in ((5 * e, False), (math.nextafter(5 * e, 0.0), True)):
            ref, dev = both(ctx, g, 1, x, z, c, mincoeff=mc, batched=batched)
            assert_same_terms(dev, ref, exact=True, what=f"mincoeff={mc!r}")
            has = any(int(dev.x[i, 0]) == 0 and int(dev.z[i, 0]) == 1 for i in range(len(dev)))
            assert has == kept, (len(g), mc)
