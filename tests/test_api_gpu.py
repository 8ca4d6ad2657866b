"""The drop-in `propagate` (fastfermion_amd._ffcore -> libffamd.so) on the GPU,
written like the reference's own suites:

  test/pauli_circuit_test.py:53-111   propagate vs two naive oracles (aspoly
                                      products and gate calls), 1e-14
  test/omp_test.py, sharded_test.py   same circuits; here the GPU result is
                                      compared with the CPU oracle's serial run
  test/batched_test.py                schedule equivalence
  test/majorana_circuit_test.py       Majorana propagate vs naive oracles
  examples/pauli_propagation.py       su4circuit-16 golden <Z15>
"""
from __future__ import annotations

import os
import warnings

import numpy as np
import pytest

import pyffo
from fastfermion_amd import workloads as wl
import fastfermion_amd as ff
from fastfermion_amd import CNOT, H, MROT, ROT, S, MajoranaString, propagate

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def approx_equal(a, b, tol=1e-14):
    return (a - b).norm("inf") <= tol


def simple_propagate_1(circuit, obs):
    for gate in reversed(circuit):
        gp = gate.aspoly()
        obs = (gp.dagger() * obs * gp).compress(0)
    return obs


def simple_propagate_2(circuit, obs):
    for gate in reversed(circuit):
        obs = gate(obs).compress(0)
    return obs


def test_propagate_vs_naive_oracles():
    circuits = [
        [H(0)], [S(0)], [CNOT(0, 1)], [H(0), CNOT(0, 1), S(1), CNOT(1, 2), H(2)],
        [ROT("I", 0.3)], [ROT("XX", 0.34)],
        [H(0), CNOT(0, 1), ROT("XY", 0.23), ROT("ZZ", 0.78), S(1), CNOT(1, 2), H(2)],
    ]
    observables = [ff.poly("X0 Z1"), ff.poly("X0Z1Y2"), ff.poly("-Z2+3*X0")]
    for circuit in circuits:
        for obs in observables:
            res = propagate(circuit, obs)
            assert approx_equal(simple_propagate_1(circuit, obs), res.compress(0))
            assert approx_equal(simple_propagate_2(circuit, obs), res.compress(0))


def to_oracle(p, W=1):
    x, z, c = p.to_arrays()
    return x[:, :W].copy(), z[:, :W].copy(), c


def tfim(n, dt):
    return [ROT("ZZ", [i, (i + 1) % n], -2 * dt) for i in range(n)] + [ROT("X", [i], -2 * dt) for i in range(n)]


def heisenberg(n, dt):
    g = []
    for i in range(n):
        j = (i + 1) % n
        g += [ROT("XX", [i, j], 2 * dt), ROT("YY", [i, j], 2 * dt), ROT("ZZ", [i, j], 2 * dt)]
    return g


def pack(circuit):
    """ff gate objects -> packed gate table (for the oracle)."""
    from fastfermion_amd import _abi
    recs = []
    for g in circuit:
        if isinstance(g, ROT):
            r = np.zeros(1, _abi.PAULI_GATE_DTYPE)
            r["kind"] = _abi.ROT
            r["theta"] = g.theta
            x, z, _ = ff.PauliPolynomial(g.axis).to_arrays()
            r["x"][0] = x[0]
            r["z"][0] = z[0]
            recs.append(r)
        else:
            kind = {"H": _abi.H, "S": _abi.S, "CNOT": _abi.CNOT, "SWAP": _abi.SWAP, "CZ": _abi.CZ}[type(g).__name__]
            q = g.qubits
            recs.append(_abi.clifford_record(kind, q[0], q[1] if len(q) > 1 else 0))
    return np.concatenate(recs) if recs else np.zeros(0, _abi.PAULI_GATE_DTYPE)


def assert_matches_oracle(circuit, obs, **kw):
    res = propagate(circuit, obs, **kw)
    o = obs if isinstance(obs, ff.PauliPolynomial) else ff.PauliPolynomial(obs)
    x, z, c = to_oracle(o)
    okw = {k: v for k, v in kw.items() if k in ("maxdegree", "mincoeff", "batched", "topk", "max_xweight", "xtrunc_period")}
    if okw.get("maxdegree") is None:
        okw.pop("maxdegree", None)
    if okw.get("mincoeff") is None:
        okw.pop("mincoeff", None)
    ref = pyffo.pauli_propagate(pack(circuit), 1, x, z, c, **okw)
    assert len(res) == len(ref)
    rd = {(int(ref.x[i, 0]), int(ref.z[i, 0])): ref.c[i] for i in range(len(ref.c))}
    rx, rz, rc = res.to_arrays()
    for i in range(len(rc)):
        assert rc[i] == rd[(int(rx[i, 0]), int(rz[i, 0]))]
    assert abs(res.overlapwithzero() - pyffo.pauli_expect(ref, [0])) < 1e-12
    return res


@pytest.mark.parametrize("kw", [{}, {"mincoeff": 1e-3}, {"maxdegree": 3}, {"batched": False, "mincoeff": 1e-3}])
def test_tfim_suite(kw):
    assert_matches_oracle(tfim(6, 0.1) * 5, ff.PauliString("Z" + "I" * 5), **kw)


def test_heisenberg_suite():
    assert_matches_oracle(heisenberg(4, 0.03) * 5, ff.PauliString("ZIII"))


def test_mixed_clifford_rot():
    c = [H(0), ROT("Z", (0,), 0.5), CNOT(0, 1), ROT("XX", [0, 1], 0.3)]
    for batched in (True, False):
        assert_matches_oracle(c * 5, ff.PauliString("ZI"), batched=batched)


def test_polynomial_observable_and_edge_cases():
    obs = ff.PauliPolynomial(ff.PauliString("ZI")) + ff.PauliPolynomial(ff.PauliString("IZ"))
    assert_matches_oracle(tfim(4, 0.05) * 10, obs)
    assert_matches_oracle([], ff.PauliString("Z"))
    assert_matches_oracle([ROT("ZZ", [0, 1], 0.3)] * 10, ff.PauliString("ZI"))
    assert_matches_oracle([H(0), CNOT(0, 1)], ff.PauliString("ZI"))


def test_identity_rotation_and_thread_warning():
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        res = propagate([ROT("X", (0,), 0.0)] * 10, ff.PauliString("Z"), n_threads=4)
    assert any(issubclass(x.category, RuntimeWarning) for x in w)
    assert abs(res.overlapwithzero() - 1.0) < 1e-14


def test_batched_equals_unbatched_without_truncation():
    for circuit, steps, obs in ((tfim(6, 0.1), 5, "ZIIIII"), (heisenberg(4, 0.03), 5, "ZIII"),
                                ([ROT("X", [i], 0.1) for i in range(6)], 3, "ZIIIII")):
        a = propagate(circuit * steps, ff.PauliString(obs), batched=False)
        b = propagate(circuit * steps, ff.PauliString(obs), batched=True)
        assert len(a) == len(b) and approx_equal(a, b, 1e-12)
    a = propagate(tfim(6, 0.1) * 5, ff.PauliString("ZIIIII"), batched=False, mincoeff=1e-3)
    b = propagate(tfim(6, 0.1) * 5, ff.PauliString("ZIIIII"), batched=True, mincoeff=1e-3)
    assert abs(a.overlapwithzero() - b.overlapwithzero()) < 1e-3


def test_su4_golden():
    g = wl.read_circuit(os.path.join(GOLDEN, "su4circuit-16.txt"))
    circuit = []
    lines = open(os.path.join(GOLDEN, "su4circuit-16.txt")).read().split("\n")[1:]
    for line in lines:
        p = line.split()
        if p:
            circuit.append(ROT(p[1], [int(q) for q in p[2:-1]], float(p[-1])))
    assert len(circuit) == len(g) == 1440
    res = propagate(circuit, ff.poly("Z15"), maxdegree=4)
    assert abs(res.overlapwithzero() - 0.008664) <= 1e-6
    e = ff.propagate_expectation(circuit, ff.poly("Z15"), maxdegree=4)
    assert abs(e - res.overlapwithzero()) < 1e-12


def test_precision_variants_through_api():
    circuit = tfim(8, 0.07) * 6
    a = propagate(circuit, ff.PauliString("Z0"), mincoeff=1e-6)
    for prec, tol in (("fp32c", 1e-12), ("fp32", 1e-5)):
        b = propagate(circuit, ff.PauliString("Z0"), mincoeff=1e-6, precision=prec)
        assert abs(a.overlapwithzero() - b.overlapwithzero()) < 10 * tol


def test_majorana_propagate_vs_naive():
    circuits = [[MROT(MajoranaString([0, 2]), 0.2j, 0.1), MROT(MajoranaString([0, 1, 2, 3]), 0.3, 0.1)]]
    observables = [ff.poly("m0 m1"), ff.poly("-m1 + m0m2"), ff.poly("-1j*m2 + 2.1 m1 m0")]
    for circuit in circuits:
        for obs in observables:
            res = propagate(circuit, obs)
            assert (simple_propagate_1(circuit, obs) - res.compress(0)).norm("inf") <= 1e-14
            assert (simple_propagate_2(circuit, obs) - res.compress(0)).norm("inf") <= 1e-14


def test_majorana_truncation_through_api():
    rng = np.random.default_rng(1)
    circ = []
    for _ in range(40):
        d = int(rng.integers(1, 5))
        S = sorted(rng.choice(16, d, replace=False).tolist())
        circ.append(MROT(MajoranaString(S), float(rng.uniform(-1, 1))))
    obs = ff.poly("0.5 + 0.5j m0 m1")
    res = propagate(circ, obs, maxdegree=6, mincoeff=1e-5)
    gates = wl.pack_majorana([(g.axis.indices(), g.theta, (g.axis.degree() >> 1) & 1) for g in circ])
    m = np.array([[0], [3]], np.uint64)
    ref = pyffo.majorana_propagate(gates, 1, m, np.array([0.5, 0.5j]), maxdegree=6, mincoeff=1e-5)
    assert len(res) == len(ref)
    for occ in ([], [0], [1, 3]):
        want = pyffo.majorana_expect(ref, [sum(1 << q for q in occ)])
        assert abs(ff.FockState(occ)(res) - want) < 1e-12


def test_profile_dictionary_has_reference_keys():
    """profile_reset / profile_get (src/python_bind.cpp:427-446): the reference's
    keys are present, plus the device counters."""
    ff.profile_reset()
    circ = tfim(6, 0.1) * 3
    propagate(circ, ff.PauliString("Z" + "I" * 5))
    d = ff.profile_get()
    for k in ("snapshot", "emit", "merge", "n_gates", "emit_per_thread", "merge_per_thread", "n_local", "n_remote",
              "updates", "peak_terms"):
        assert k in d, k
    assert d["n_gates"] == len(circ)
    assert d["updates"] > 0 and d["emit"] >= 0.0
    ff.profile_disable()


def test_negative_maxdegree_is_literal():
    """maxdegree compares literally in both overloads (src/pauli/propagate.h:82,
    src/majorana_propagation.h:43, bound at src/python_bind.cpp:404, :505-511):
    -1 keeps no partner, so one rotation leaves only the cos-scaled terms."""
    import math
    t = 0.37
    res = propagate([ROT("X", [0], t)], ff.poly("Z0 + 0.5 Z1"), maxdegree=-1)
    assert len(res) == 2
    assert abs(res.coefficient(ff.PauliString("ZI")) - math.cos(t)) < 1e-15
    assert len(propagate([ROT("X", [0], t)], ff.poly("Z0 + 0.5 Z1"))) == 3
    circ = [MROT(MajoranaString([0, 1]), 1j, t)]
    mres = propagate(circ, ff.poly("m1 m2"), maxdegree=-1)
    assert len(mres) == 1
    assert len(propagate(circ, ff.poly("m1 m2"))) == 2
