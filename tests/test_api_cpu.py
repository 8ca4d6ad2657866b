"""Host-side logic of the drop-in module (no GPU): operator types, gates,
states, the poly() parser, and loud failure without a device.

Mirrors the reference checks that need no propagation: the Clifford gates on
all 3-qubit strings (test/pauli_circuit_test.py:8-51, cirq replaced by dense
matrices), MROT against scipy expm (test/majorana_circuit_test.py:43-74),
product-state readouts against dense vectors (test/state_test.py:6-70).
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

import _dense as D
import fastfermion_amd as ff


def pp_dense(p, n):
    M = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for ps, c in p.terms.items():
        M += c * D.pauli_matrix(ps.to_string(n))
    return M


def mp_dense(p, modes):
    M = np.zeros((2 ** modes, 2 ** modes), dtype=complex)
    for ms, c in p.terms.items():
        M += c * D.majorana_matrix(ms.indices(), modes)
    return M


def test_pauli_string_forms_and_errors():
    assert ff.PauliString("XXIYZ").indices() == [(0, "X"), (1, "X"), (3, "Y"), (4, "Z")]
    assert ff.PauliString("X0 Y1 Z4") == ff.PauliString([(0, "X"), (1, "Y"), (4, "Z")])
    assert ff.PauliString("XIZ").to_string(5) == "XIZII"
    assert ff.PauliString("X3 Y7").extent() == 8
    assert ff.PauliString("XYZ").degree() == 3 and ff.PauliString("XYZ").degree("Y") == 1
    with pytest.raises(RuntimeError):
        ff.PauliString([(0, "X"), (0, "Z")])
    with pytest.raises(RuntimeError):
        ff.PauliString([(0, "Q")])
    with pytest.raises(RuntimeError):
        ff.PauliString([(ff.MAX_QUBITS, "X")])
    assert ff.PauliString([(ff.MAX_QUBITS - 1, "Y")]).extent() == ff.MAX_QUBITS


def test_pauli_products_exhaustive_two_qubits():
    for a in itertools.product("IXYZ", repeat=2):
        for b in itertools.product("IXYZ", repeat=2):
            a_, b_ = "".join(a), "".join(b)
            p = ff.PauliString(a_) * ff.PauliString(b_)
            assert np.allclose(pp_dense(p, 2), D.pauli_matrix(a_) @ D.pauli_matrix(b_), atol=1e-15)
            assert ff.PauliString(a_).commutes(ff.PauliString(b_)) == np.allclose(
                D.pauli_matrix(a_) @ D.pauli_matrix(b_), D.pauli_matrix(b_) @ D.pauli_matrix(a_))


def test_polynomial_arithmetic():
    p = ff.poly("0.5 X0 Z1 + 1j Y2 - 3")
    q = ff.poly("Z1 + 2 X0")
    n = 3
    assert np.allclose(pp_dense(p * q, n), pp_dense(p, n) @ pp_dense(q, n))
    assert np.allclose(pp_dense(p + q, n), pp_dense(p, n) + pp_dense(q, n))
    assert np.allclose(pp_dense(p - 2 * q, n), pp_dense(p, n) - 2 * pp_dense(q, n))
    assert np.allclose(pp_dense(p.dagger(), n), pp_dense(p, n).conj().T)
    assert (p - p) == ff.PauliPolynomial() and len(p - p) == 0
    assert p.coefficient("X0 Z1") == 0.5 and p.coefficient([(2, "Y")]) == 1j
    assert abs(p.norm(2) - math.sqrt(0.25 + 1 + 9)) < 1e-15
    assert ff.PauliPolynomial([(0, "X"), (0, "Y")], 2.0) == ff.poly("2j Z0")
    assert ff.poly("3.1 (X1 + Y1) Z1") == ff.poly("-3.1j Y1 + 3.1j X1")


def test_clifford_gates_all_three_qubit_strings():
    """gate(P) = U^dag P U for H, S, CNOT, SWAP, CZ on every placement."""
    n = 3
    gates = {"H": (ff.H, 1, 1), "S": (ff.S, 1, 2), "CNOT": (ff.CNOT, 2, 3), "SWAP": (ff.SWAP, 2, 4), "CZ": (ff.CZ, 2, 5)}
    for name, (cls, nq, kind) in gates.items():
        for qs in itertools.permutations(range(n), nq):
            g = cls(*qs)
            U = D.CLIFFORD[kind](n, qs[0], qs[1] if nq > 1 else 0)
            for s in itertools.product("IXYZ", repeat=n):
                s = "".join(s)
                out, c = g(ff.PauliString(s))
                assert np.allclose(c * D.pauli_matrix(out.to_string(n)), U.conj().T @ D.pauli_matrix(s) @ U), \
                    (name, qs, s)
            assert np.allclose(pp_dense(g.aspoly(), n), U)


def test_rot_gate_matches_dense():
    n = 3
    for axis in ("XX", "YZ", "ZIY", "Y"):
        g = ff.ROT(axis, 0.37)
        U = D.rot_unitary(axis + "I" * (n - len(axis)), 0.37)
        assert np.allclose(pp_dense(g.aspoly(), n), U)
        A = ff.poly("0.3 X0 + Z1 Y2 - 0.5j Z0 Z1")
        assert np.allclose(pp_dense(g(A), n), U.conj().T @ pp_dense(A, n) @ U, atol=1e-14)
    g = ff.ROT("ZZ", [3, 5], 1.1)
    assert g.qubits == [3, 5] and g.theta == 1.1 and g.axis == ff.PauliString("Z3 Z5")


def test_mrot_matches_expm():
    """test/majorana_circuit_test.py:43-74 with Jordan-Wigner dense matrices."""
    modes = 2
    A = ff.poly("0.1 m0 m1 + 1j m0 - 4 m1")
    for S, t in (([], 0.3), ([1], 0.3), ([0, 1], 0.3), ([0, 1, 3], 0.2), ([0, 1, 2, 3], 0.2)):
        U = ff.MROT(ff.MajoranaString(S), t)
        Up = U.aspoly()
        assert (Up.dagger() * A * Up - U(A)).norm("inf") <= 1e-14
        r = (U.axis.degree() // 2) % 2
        Vm = D.expm(-1j * U.theta / 2 * (1j) ** r * D.majorana_matrix(S, modes))
        assert np.linalg.norm(mp_dense(Up, modes) - Vm) <= 1e-14
    U = ff.MROT(ff.MajoranaString([0, 1]), 0.2j, 0.3)
    Vm = D.expm(-1j * 0.3 / 2 * 0.2j * D.majorana_matrix([0, 1], modes))
    assert np.linalg.norm(mp_dense(U.aspoly(), modes) - Vm) <= 1e-14
    with pytest.raises(RuntimeError, match="not Hermitian"):
        ff.MROT(ff.MajoranaString([0, 2]), 1.0, 2.0)


def test_majorana_products_and_ordering():
    modes = 3
    for a in itertools.combinations(range(6), 2):
        for b in itertools.combinations(range(6), 3):
            p = ff.MajoranaString(list(a)) * ff.MajoranaString(list(b))
            assert np.allclose(mp_dense(p, modes), D.majorana_matrix(a, modes) @ D.majorana_matrix(b, modes))
    assert ff.MajoranaPolynomial([1, 3, 2]) == -1 * ff.MajoranaPolynomial([1, 2, 3])
    with pytest.raises(RuntimeError):
        ff.MajoranaString([2, 1])


def test_product_states_vs_dense():
    """test/state_test.py:6-70 for the Pauli and Majorana readouts."""
    assert ff.QubitProductState().vec(1) == [1, 0]
    assert ff.QubitProductState([0]).vec(2) == [0, 0, 1, 0]
    assert ff.QubitProductState([1]).vec(2) == [0, 1, 0, 0]
    assert ff.FockState([0, 1]).vec(2) == [0, 0, 0, 1]
    n = 4
    for s in itertools.product("IXYZ", repeat=n):
        s = "".join(s)
        ps = ff.PauliString(s)
        M = D.pauli_matrix(s)
        for B in itertools.chain.from_iterable(itertools.combinations(range(n), r) for r in range(n + 1)):
            st = ff.QubitProductState(list(B))
            v = np.array(st.vec(n))
            assert abs(st(ps) - v @ M @ v) <= 1e-14
    m = 3
    for r in range(2 * m + 1):
        for S in itertools.combinations(range(2 * m), r):
            ms = ff.MajoranaString(list(S))
            M = D.majorana_matrix(S, m)
            for B in itertools.chain.from_iterable(itertools.combinations(range(m), k) for k in range(m + 1)):
                st = ff.FockState(list(B))
                v = np.array(st.vec(m))
                assert abs(st(ms) - v @ M @ v) <= 1e-14


def test_overlapwithzero():
    p = ff.poly("0.5 Z0 Z1 + 2 X0 - 0.25 Z3 + 1")
    assert p.overlapwithzero() == 1.25
    assert ff.QubitProductState([0])(p) == 1 - 0.5 - 0.25


def test_arrays_roundtrip():
    p = ff.poly("0.5 Z0 Z1 + 2j X0 - 0.25 Z200")
    x, z, c = p.to_arrays()
    q = ff.PauliPolynomial.from_arrays(x, z, c)
    assert q == p and len(q) == 3


def test_propagate_fails_loudly_without_device():
    if ff.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError, match="no usable HIP device"):
        ff.propagate([ff.ROT("X", [0], 0.3)], ff.PauliString("Z0"))
    with pytest.raises(RuntimeError, match="no usable HIP device"):
        ff.propagate([ff.MROT(ff.MajoranaString([0, 1]), 0.3)], ff.MajoranaString([0]))


def test_jordan_wigner_bridge_vs_dense():
    """jw / tomajorana (src/transforms.h:125-160, 254-298) against dense
    Jordan-Wigner matrices on 3 modes, every Majorana string and every Pauli string."""
    n = 3
    for r in range(2 * n + 1):
        for S in itertools.combinations(range(2 * n), r):
            p = ff.jw(ff.MajoranaString(list(S)))
            assert np.allclose(pp_dense(p, n), D.majorana_matrix(S, n), atol=1e-14), S
            back = ff.tomajorana(p)
            assert back == ff.MajoranaPolynomial(ff.MajoranaString(list(S))), S
    for s in itertools.product("IXYZ", repeat=n):
        s = "".join(s)
        m = ff.PauliString(s).tomajorana()
        assert np.allclose(mp_dense(m, n), D.pauli_matrix(s), atol=1e-14), s
    A = ff.poly("0.3 m0 m3 + 1j m1 - 0.5 m0 m1 m2 m5")
    assert np.allclose(pp_dense(ff.jw(A), n), mp_dense(A, n), atol=1e-14)
    assert ff.topauli(A) == ff.jw(A) == A.topauli()


def test_jordan_wigner_rotation():
    """The JW image of MROT is the ROT with the same unitary."""
    n = 3
    for S, c, t in (([0, 1], 1.0j, 0.3), ([1, 4], -0.5j, 0.7), ([0, 1, 2, 3], -1.0, 0.2), ([2, 5], 2.0j, -0.4),
                    ([0, 3, 4, 5], 0.25, 1.1)):
        g = ff.MROT(ff.MajoranaString(S), c, t)
        r = ff.jw(g)
        assert np.allclose(pp_dense(r.aspoly(), n), mp_dense(g.aspoly(), n), atol=1e-14), S


def test_fermi_hubbard_terms_vs_fermionic_hamiltonian():
    """The Majorana form of the Hubbard Hamiltonian equals the one built from
    dense fermionic operators a_j = (m_2j + i m_2j+1)/2 (2x1 lattice, 4 modes),
    up to the dropped constant U/4 per site."""
    from fastfermion_amd import workloads as wl
    lx, ly, t, u = 2, 1, 0.7, 2.3
    nm = 2 * lx * ly
    a = [(D.majorana_matrix([2 * j], nm) + 1j * D.majorana_matrix([2 * j + 1], nm)) / 2 for j in range(nm)]
    H = np.zeros((2 ** nm, 2 ** nm), complex)
    for s1, s2 in ((0, 1),):
        for spin in (0, 1):
            i, j = 2 * s1 + spin, 2 * s2 + spin
            H += -t * (a[i].conj().T @ a[j] + a[j].conj().T @ a[i])
    for s in range(lx * ly):
        nu = a[2 * s].conj().T @ a[2 * s]
        nd = a[2 * s + 1].conj().T @ a[2 * s + 1]
        H += u * nu @ nd - u / 4 * np.eye(2 ** nm)
    M = np.zeros_like(H)
    for idx, c in wl.fermi_hubbard_terms(lx, ly, t, u):
        M += c * D.majorana_matrix(idx, nm)
    assert np.allclose(H, M, atol=1e-13)


@pytest.mark.parametrize("lx,ly,pbc", [(6, 6, False), (6, 6, True), (5, 5, True), (3, 4, False), (2, 2, False)])
def test_fermi_hubbard_layered_order(lx, ly, pbc):
    """The layered Trotter order holds the same monomials as the bond order, in
    groups with pairwise disjoint index sets: four bond matchings (a fifth and
    sixth for wrap-around bonds of odd sides), the degree-2 on-site monomials,
    then the degree-4 ones."""
    from fastfermion_amd import workloads as wl
    a = wl.fermi_hubbard_terms(lx, ly, pbc=pbc)
    b = wl.fermi_hubbard_terms(lx, ly, pbc=pbc, order="layers")
    assert sorted((tuple(i), c) for i, c in a) == sorted((tuple(i), c) for i, c in b)
    groups, used = 1, set()
    for idx, _ in b:
        if used & set(idx):
            groups += 1
            used = set()
        used |= set(idx)
    n_match = sum(1 for side in (lx, ly) for m in range(3)
                  if m < 2 and side > 1 + m or (m == 2 and pbc and side > 2 and side % 2))
    assert groups <= n_match + 2
    # the last two groups: every on-site degree-2 monomial, then every degree-4 one
    n = lx * ly
    assert [len(i) for i, _ in b[-3 * n:]] == [2] * (2 * n) + [4] * n
    with pytest.raises(ValueError):
        wl.fermi_hubbard_terms(lx, ly, order="rows")
