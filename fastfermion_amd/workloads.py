"""Synthetic circuits of the benchmark and test workloads, as packed gate tables.

Every random draw comes from SplitMix64 with u = (x >> 11) * 2^-53, so the C++,
Python and HIP sides can regenerate identical inputs from a seed.
"""
from __future__ import annotations

import math

import numpy as np

from ._abi import MAJORANA_GATE_DTYPE, PAULI_GATE_DTYPE, ROT, pauli_word, words_for_bits


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53


def _pack(gates):
    out = np.zeros(len(gates), dtype=PAULI_GATE_DTYPE)
    for i, (ops, qubits, theta) in enumerate(gates):
        out[i]["kind"] = ROT
        out[i]["theta"] = theta
        x, z = pauli_word(ops, qubits)
        out[i]["x"] = x
        out[i]["z"] = z
    return out


def tfim_chain(n: int = 20, steps: int = 5, zz: float = 0.2, rx: float = 0.1):
    """Open-chain transverse-field Ising Trotter circuit: per step RZZ(zz) on the
    n-1 bonds (i, i+1), then RX(rx) on every qubit."""
    g = []
    for _ in range(steps):
        g += [("ZZ", (i, i + 1), zz) for i in range(n - 1)]
        g += [("X", (i,), rx) for i in range(n)]
    return _pack(g)


def square_edges(L: int):
    """Open L x L lattice, site s = L*r + c: horizontal bonds (s, s+1) in
    row-major order, then vertical bonds (s, s+L)."""
    h = [(L * r + c, L * r + c + 1) for r in range(L) for c in range(L - 1)]
    v = [(s, s + L) for s in range(L * (L - 1))]
    return h + v


def kicked_ising(L: int = 10, layers: int = 20, theta_lo: float = 0.2, theta_hi: float = 0.6,
                 seed: int = 2026, zz: float = -math.pi / 2, thetas=None):
    """Square-lattice kicked Ising: per layer RZZ(zz) on every edge, then RX(theta_s)
    on every site; theta_s = lo + (hi-lo)*u_s drawn once per site (seeded)."""
    n = L * L
    if thetas is None:
        rng = SplitMix64(seed)
        thetas = [theta_lo + (theta_hi - theta_lo) * rng.uniform() for _ in range(n)]
    edges = square_edges(L)
    g = []
    for _ in range(layers):
        g += [("ZZ", e, zz) for e in edges]
        g += [("X", (s,), thetas[s]) for s in range(n)]
    return _pack(g), list(thetas)


def single_pauli(ops: str, qubits, nqubits: int, coeff: complex = 1.0):
    """Observable with one Pauli term, as term-major key arrays."""
    W = words_for_bits(nqubits)
    x, z = pauli_word(ops, qubits, W)
    return x.reshape(1, W), z.reshape(1, W), np.array([coeff], dtype=np.complex128)


def pack_majorana(gates):
    """gates: iterable of (indices, theta, r) -> packed Majorana gate table."""
    out = np.zeros(len(gates), dtype=MAJORANA_GATE_DTYPE)
    for i, (idx, theta, r) in enumerate(gates):
        m = np.zeros(4, dtype=np.uint64)
        for a in idx:
            m[a >> 6] |= np.uint64(1) << np.uint64(a & 63)
        out[i]["m"] = m
        out[i]["theta"] = theta
        out[i]["r"] = r
    return out


def read_circuit(path: str):
    """Gate table from the text circuit format of fastfermion's examples
    (examples/common.py:3-35): first line the gate count, then one gate per
    line: `ROT <axis letters> <qubits...> <theta>`, or `H i`, `S i`, `CNOT i j`,
    `SWAP i j`, `CZ i j`."""
    from ._abi import CNOT, CZ, H, S, SWAP, clifford_record, rot_record

    kinds = {"H": H, "S": S, "CNOT": CNOT, "SWAP": SWAP, "CZ": CZ}
    recs = []
    with open(path, "r", encoding="utf-8") as f:
        n = int(f.readline())
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "ROT":
                recs.append(rot_record(parts[1], [int(q) for q in parts[2:-1]], float(parts[-1])))
            else:
                qs = [int(q) for q in parts[1:]]
                recs.append(clifford_record(kinds[parts[0]], qs[0], qs[1] if len(qs) > 1 else 0))
    if len(recs) != n:
        raise ValueError(f"{path}: header says {n} gates, found {len(recs)}")
    return np.concatenate(recs) if recs else np.zeros(0, PAULI_GATE_DTYPE)


def batch_kicked_ising(n_inst: int = 512, seed: int = 4242, l_lo: int = 6, l_span: int = 7, d_lo: int = 4,
                       d_span: int = 21, zz: float = -math.pi / 2):
    """Uneven kicked-Ising instances (the batched configuration): instance i has
    an L x L lattice with L = l_lo + (u mod l_span), depth d_lo + (u mod d_span)
    layers, per-qubit RX angles uniform in [0, pi/2], and a random weight-1 or
    weight-2 Z observable, all drawn from one SplitMix64 stream.  Returns a list
    of (gates, x, z, c, L, depth) with keys packed on words_for_bits(L*L) words."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n_inst):
        L = l_lo + int(rng.next() % l_span)
        depth = d_lo + int(rng.next() % d_span)
        nq = L * L
        thetas = [0.5 * math.pi * rng.uniform() for _ in range(nq)]
        gates, _ = kicked_ising(L, depth, zz=zz, thetas=thetas)
        w = 1 + int(rng.next() % 2)
        sites = []
        while len(sites) < w:
            q = int(rng.next() % nq)
            if q not in sites:
                sites.append(q)
        x, z, c = single_pauli("Z" * w, sorted(sites), nq)
        out.append((gates, x, z, c, L, depth))
    return out


def fermi_hubbard_terms(lx: int, ly: int, t: float = 1.0, u: float = 4.0, pbc: bool = False,
                        order: str = "bonds"):
    """Fermi-Hubbard Hamiltonian -t sum_<ij>,s (a_is^dag a_js + h.c.) + U sum_i n_i,up n_i,dn
    on an lx x ly lattice as Majorana monomials [(indices, coefficient)], in the
    Trotter order used by fermi_hubbard (the constant U/4 per site is dropped).
    Site s = x + lx*y, mode 2s + spin, Majoranas 2*mode and 2*mode + 1 with
    a_j = (m_2j + i m_2j+1)/2, so n_j = (1 + i m_2j m_2j+1)/2 and
    a_i^dag a_j + h.c. = (i/2)(m_2i m_2j+1 - m_2i+1 m_2j).

    order "bonds": horizontal bonds row by row, then vertical ones (both spins,
    both hopping monomials per bond), then per site the two degree-2 on-site
    monomials and the degree-4 one.  order "layers": the bonds in four
    matchings (horizontal bonds at even x, at odd x, vertical bonds at even y,
    at odd y; a wrap-around bond that meets another bond of its matching gets a
    matching of its own), then every degree-2 on-site monomial, then every
    degree-4 one: each group's monomials have pairwise disjoint index sets.
    With pbc also the wrap-around bonds (lattice sides above 2)."""
    if order not in ("bonds", "layers"):
        raise ValueError(f"unknown Trotter order {order!r}")

    def site(x, y):
        return x + lx * y

    hb, vb = [], []  # (matching, s1, s2)
    for y in range(ly):
        for x in range(lx - 1 + (1 if pbc and lx > 2 else 0)):
            wrap = x == lx - 1
            hb.append((2 if wrap and lx % 2 else x % 2, site(x, y), site((x + 1) % lx, y)))
    for y in range(ly - 1 + (1 if pbc and ly > 2 else 0)):
        for x in range(lx):
            wrap = y == ly - 1
            vb.append((2 if wrap and ly % 2 else y % 2, site(x, y), site(x, (y + 1) % ly)))
    if order == "bonds":
        bonds = [(s1, s2) for _, s1, s2 in hb + vb]
    else:
        bonds = [(s1, s2) for grp in (hb, vb) for m in range(3) for mm, s1, s2 in grp if mm == m]

    def hop(s1, s2):
        a, b = min(s1, s2), max(s1, s2)
        out = []
        for spin in (0, 1):
            i, j = 2 * a + spin, 2 * b + spin
            out.append(([2 * i, 2 * j + 1], -0.5j * t))
            out.append(([2 * i + 1, 2 * j], 0.5j * t))
        return out

    def onsite(s):
        up, dn = 2 * s, 2 * s + 1
        return [([2 * up, 2 * up + 1], 0.25j * u), ([2 * dn, 2 * dn + 1], 0.25j * u),
                ([2 * up, 2 * up + 1, 2 * dn, 2 * dn + 1], -0.25 * u)]

    terms = [m for (s1, s2) in bonds for m in hop(s1, s2)]
    sites = [onsite(s) for s in range(lx * ly)]
    if order == "bonds":
        terms += [m for ms in sites for m in ms]
    else:
        terms += [m for ms in sites for m in ms[:2]] + [ms[2] for ms in sites]
    return terms


def fermi_hubbard(lx: int = 6, ly: int = 6, t: float = 1.0, u: float = 4.0, dt: float = 0.05, steps: int = 20,
                  pbc: bool = False, order: str = "bonds"):
    """First-order Trotter circuit of fermi_hubbard_terms: per step every
    monomial c*M in order as MROT(M, c, 2 dt), i.e. theta = Im(c) 2dt for
    degree 2 (r = 1) and Re(c) 2dt for degree 4 (r = 0).  Returns (packed gate
    table, Majorana observable n_{0,up} = (1 + i m0 m1)/2 as (m, c))."""
    step = []
    for idx, c in fermi_hubbard_terms(lx, ly, t, u, pbc, order):
        r = (len(idx) >> 1) & 1
        step.append((idx, (c.imag if r else c.real) * 2 * dt, r))
    gates = pack_majorana(step * steps)
    nmaj = 4 * lx * ly
    W = words_for_bits(nmaj)
    m = np.zeros((2, W), np.uint64)
    m[1, 0] = np.uint64(0b11)
    c = np.array([0.5, 0.5j])
    return gates, (m, c)


def lattice_geometry(L: int):
    """Rows and columns of the qubits of an L x L lattice (site s = L*r + c)."""
    return [s // L for s in range(L * L)], [s % L for s in range(L * L)]


def hubbard_geometry(lx: int, ly: int):
    """Rows and columns of every Majorana index of fermi_hubbard_terms: Majorana
    m belongs to mode m // 2, site m // 4 = x + lx*y."""
    n = 4 * lx * ly
    return [(m // 4) // lx for m in range(n)], [(m // 4) % lx for m in range(n)]


def kicked_ising_family(L: int = 10, layers: int = 30, theta_lo: float = 0.3, theta_hi: float = 0.7, seed: int = 7):
    """Kicked-Ising circuit of the learned-pruning workload: RZZ(-pi/2) on every
    bond, then RX(theta_s) with theta_s uniform in [theta_lo, theta_hi] (seeded)."""
    return kicked_ising(L, layers, theta_lo=theta_lo, theta_hi=theta_hi, seed=seed)
