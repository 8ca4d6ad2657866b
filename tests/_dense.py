"""Dense-matrix brute force used to pin the oracle (small systems only).

Qubit q of an n-qubit register is the q-th most significant tensor factor, so
a computational basis state with qubit set {q} has index sum 2^(n-1-q): the
convention of QubitProductState.vec.  Majorana operators use the Jordan-Wigner
image m_{2j} = Z...Z X_j, m_{2j+1} = Z...Z Y_j (Z on modes < j).
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.linalg

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
LETTER = {"I": I2, "X": X, "Y": Y, "Z": Z}


def pauli_matrix(letters: str) -> np.ndarray:
    m = np.ones((1, 1), dtype=complex)
    for ch in letters:
        m = np.kron(m, LETTER[ch])
    return m


def letters_from_words(x: int, z: int, n: int) -> str:
    out = []
    for q in range(n):
        a, b = (x >> q) & 1, (z >> q) & 1
        out.append("Y" if a and b else "X" if a else "Z" if b else "I")
    return "".join(out)


def words_from_letters(s: str):
    x = z = 0
    for q, ch in enumerate(s):
        if ch in "XY":
            x |= 1 << q
        if ch in "YZ":
            z |= 1 << q
    return x, z


def pauli_decompose(M: np.ndarray, n: int, tol=1e-13):
    """{(x, z): coefficient} with M = sum c P (Tr(P M)/2^n)."""
    out = {}
    for letters in itertools.product("IXYZ", repeat=n):
        s = "".join(letters)
        c = np.trace(pauli_matrix(s) @ M) / 2 ** n
        if abs(c) > tol:
            out[words_from_letters(s)] = c
    return out


def rot_unitary(letters: str, theta: float) -> np.ndarray:
    """exp(-i theta/2 P)."""
    P = pauli_matrix(letters)
    return np.cos(theta / 2) * np.eye(len(P)) - 1j * np.sin(theta / 2) * P


CLIFFORD = {
    1: lambda n, i, j: _single(n, i, np.array([[1, 1], [1, -1]]) / np.sqrt(2)),
    2: lambda n, i, j: _single(n, i, np.array([[1, 0], [0, 1j]])),
    3: lambda n, i, j: _cnot(n, i, j),
    4: lambda n, i, j: _swap(n, i, j),
    5: lambda n, i, j: _cz(n, i, j),
}


def _single(n, i, u):
    m = np.ones((1, 1), dtype=complex)
    for q in range(n):
        m = np.kron(m, u if q == i else I2)
    return m


def _basis_perm(n, f):
    d = 2 ** n
    U = np.zeros((d, d), dtype=complex)
    for idx in range(d):
        bits = [(idx >> (n - 1 - q)) & 1 for q in range(n)]
        out, phase = f(bits)
        j = sum(b << (n - 1 - q) for q, b in enumerate(out))
        U[j, idx] = phase
    return U


def _cnot(n, i, j):
    def f(b):
        b = list(b)
        if b[i]:
            b[j] ^= 1
        return b, 1
    return _basis_perm(n, f)


def _swap(n, i, j):
    def f(b):
        b = list(b)
        b[i], b[j] = b[j], b[i]
        return b, 1
    return _basis_perm(n, f)


def _cz(n, i, j):
    return _basis_perm(n, lambda b: (list(b), -1 if (b[i] and b[j]) else 1))


def basis_vector(n: int, up) -> np.ndarray:
    v = np.zeros(2 ** n, dtype=complex)
    v[sum(1 << (n - 1 - q) for q in up)] = 1
    return v


def majorana_matrix(idx, modes: int) -> np.ndarray:
    """m_{i1} m_{i2} ... in the given order, Jordan-Wigner on `modes` qubits."""
    M = np.eye(2 ** modes, dtype=complex)
    for a in idx:
        j, odd = divmod(a, 2)
        letters = "Z" * j + ("Y" if odd else "X") + "I" * (modes - j - 1)
        M = M @ pauli_matrix(letters)
    return M


def majorana_decompose(M: np.ndarray, modes: int, tol=1e-13):
    """{occupation mask: coefficient} with M = sum c m_S (m_S in increasing order)."""
    out = {}
    nm = 2 * modes
    for r in range(nm + 1):
        for S in itertools.combinations(range(nm), r):
            B = majorana_matrix(S, modes)
            c = np.trace(B.conj().T @ M) / 2 ** modes
            if abs(c) > tol:
                out[sum(1 << a for a in S)] = c
    return out


def expm(A):
    return scipy.linalg.expm(A)
