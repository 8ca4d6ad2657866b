"""Record layouts of the C ABI (include/fastfermion_amd.h) as numpy dtypes.

A packed gate table is a contiguous numpy array of these records; the same bytes
are handed to ffx_pauli_propagate / ffx_majorana_propagate.
"""
from __future__ import annotations

import numpy as np

ROT, H, S, CNOT, SWAP, CZ = 0, 1, 2, 3, 4, 5
MAX_WORDS = 4

PAULI_GATE_DTYPE = np.dtype(
    [("kind", "<i4"), ("q0", "<i4"), ("q1", "<i4"), ("reserved", "<i4"),
     ("theta", "<f8"), ("x", "<u8", (MAX_WORDS,)), ("z", "<u8", (MAX_WORDS,))])
MAJORANA_GATE_DTYPE = np.dtype(
    [("m", "<u8", (MAX_WORDS,)), ("theta", "<f8"), ("r", "<i4"), ("reserved", "<i4")])
assert PAULI_GATE_DTYPE.itemsize == 88 and MAJORANA_GATE_DTYPE.itemsize == 48

FP64, FP32, FP32C = 0, 1, 2


def words_for_bits(nbits: int) -> int:
    """64-bit words needed for nbits bits (at least one)."""
    return max(1, (int(nbits) + 63) // 64)


def set_bit(words, q: int) -> None:
    words[q >> 6] |= np.uint64(1) << np.uint64(q & 63)


def pauli_word(ops, qubits, nwords: int = MAX_WORDS):
    """(x, z) word arrays for a Pauli string given as letters and qubit indices.

    Encoding I=(0,0), X=(1,0), Y=(1,1), Z=(0,1) as in the reference
    (src/pauli/algebra.h:30-33)."""
    x = np.zeros(nwords, dtype=np.uint64)
    z = np.zeros(nwords, dtype=np.uint64)
    for op, q in zip(ops, qubits):
        if op in "XY":
            set_bit(x, q)
        if op in "YZ":
            set_bit(z, q)
    return x, z


def rot_record(ops: str, qubits, theta: float):
    """One packed ROT record exp(-i theta/2 P), P given by letters on qubits."""
    g = np.zeros(1, dtype=PAULI_GATE_DTYPE)
    g["kind"] = ROT
    g["theta"] = theta
    x, z = pauli_word(ops, qubits)
    g["x"][0] = x
    g["z"][0] = z
    return g


def clifford_record(kind: int, q0: int, q1: int = 0):
    g = np.zeros(1, dtype=PAULI_GATE_DTYPE)
    g["kind"] = kind
    g["q0"] = q0
    g["q1"] = q1
    return g
