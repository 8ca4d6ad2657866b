"""ctypes front end of the CPU oracle (oracle/libffo_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product never imports it.

Gate records use the byte layout of the product C ABI (include/fastfermion_amd.h):
    Pauli gate    : int32 kind, q0, q1, reserved; float64 theta; uint64 x[4]; uint64 z[4]
    Majorana gate : uint64 m[4]; float64 theta; int32 r; int32 reserved
Keys are term-major uint64 arrays of shape (n_terms, n_words).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PAULI_GATE_DTYPE = np.dtype(
    [("kind", "<i4"), ("q0", "<i4"), ("q1", "<i4"), ("reserved", "<i4"),
     ("theta", "<f8"), ("x", "<u8", (4,)), ("z", "<u8", (4,))], align=False)
MAJORANA_GATE_DTYPE = np.dtype(
    [("m", "<u8", (4,)), ("theta", "<f8"), ("r", "<i4"), ("reserved", "<i4")], align=False)
assert PAULI_GATE_DTYPE.itemsize == 88 and MAJORANA_GATE_DTYPE.itemsize == 48

SERIAL, SHARDED, OMP = 0, 1, 2

# maxdegree value that disables the partner degree filter
NO_DEGREE_LIMIT = 0x7FFFFFFF


def _md(maxdegree):
    """None: no filter; an integer is compared literally (src/pauli/propagate.h:82)."""
    return NO_DEGREE_LIMIT if maxdegree is None else int(maxdegree)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Options(ctypes.Structure):
    _fields_ = [("maxdegree", ctypes.c_int32), ("topk", ctypes.c_int32),
                ("max_xweight", ctypes.c_int32), ("xtrunc_period", ctypes.c_int32),
                ("batched", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("mincoeff", ctypes.c_double)]


class _Terms(ctypes.Structure):
    _fields_ = [("n_terms", ctypes.c_int64), ("n_words", ctypes.c_int32),
                ("reserved", ctypes.c_int32),
                ("x", ctypes.POINTER(ctypes.c_uint64)), ("z", ctypes.POINTER(ctypes.c_uint64)),
                ("c_re", ctypes.POINTER(ctypes.c_double)), ("c_im", ctypes.POINTER(ctypes.c_double)),
                ("seconds", ctypes.c_double), ("updates", ctypes.c_int64),
                ("peak_terms", ctypes.c_int64)]


def lib(path=None):
    """Load (building first if needed) the oracle library.  `path` (first call
    only) loads another build of the same source, e.g. one with a planted mistake."""
    global _LIB
    if _LIB is None:
        path = path or os.path.join(_HERE, "libffo_oracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.ffo_pauli_propagate.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, vp, vp, vp, vp,
                                          ctypes.c_int64, ctypes.POINTER(_Options), ctypes.c_int32,
                                          ctypes.c_int32, ctypes.POINTER(_Terms)]
        L.ffo_pauli_propagate.restype = ctypes.c_int
        L.ffo_majorana_propagate.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, vp, vp, vp,
                                             ctypes.c_int64, ctypes.c_int32, ctypes.c_double,
                                             ctypes.POINTER(_Terms)]
        L.ffo_majorana_propagate.restype = ctypes.c_int
        L.ffo_pauli_expect.argtypes = [ctypes.POINTER(_Terms), vp, vp]
        L.ffo_majorana_expect.argtypes = [ctypes.POINTER(_Terms), vp, vp]
        L.ffo_free.argtypes = [ctypes.POINTER(_Terms)]
        L.ffo_rotation_counts.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.ffo_rotation_counts.restype = ctypes.c_int64
        L.ffo_max_threads.restype = ctypes.c_int32
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Result:
    """Propagated observable: term-major keys and complex coefficients."""

    def __init__(self, x, z, c, seconds, updates=0, peak_terms=0):
        self.x, self.z, self.c, self.seconds = x, z, c, seconds
        self.updates, self.peak_terms = updates, peak_terms

    def __len__(self):
        return len(self.c)

    def as_dict(self):
        W = self.x.shape[1]
        if self.z is None:
            return {tuple(int(v) for v in self.x[i]): complex(self.c[i]) for i in range(len(self.c))}
        return {(tuple(int(v) for v in self.x[i]), tuple(int(v) for v in self.z[i])): complex(self.c[i])
                for i in range(len(self.c))} if W else {}


def _collect(t: _Terms, pauli: bool) -> Result:
    n, W = t.n_terms, t.n_words
    x = np.ctypeslib.as_array(t.x, shape=(max(n, 1) * W,))[: n * W].reshape(n, W).copy()
    z = (np.ctypeslib.as_array(t.z, shape=(max(n, 1) * W,))[: n * W].reshape(n, W).copy()
         if pauli else None)
    re = np.ctypeslib.as_array(t.c_re, shape=(max(n, 1),))[:n].copy()
    im = np.ctypeslib.as_array(t.c_im, shape=(max(n, 1),))[:n].copy()
    return Result(x, z, re + 1j * im, t.seconds, t.updates, t.peak_terms)


def pauli_propagate(gates: np.ndarray, n_words: int, x: np.ndarray, z: np.ndarray, c: np.ndarray,
                    maxdegree: int | None = None, mincoeff: float = 0.0, topk: int = 0, max_xweight: int = -1,
                    xtrunc_period: int = 1, batched: bool = True, strategy: int = SERIAL,
                    n_threads: int = 1) -> Result:
    L = lib()
    gates = np.ascontiguousarray(gates, dtype=PAULI_GATE_DTYPE)
    x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, n_words)
    z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, n_words)
    c = np.asarray(c, dtype=np.complex128).reshape(-1)
    re = np.ascontiguousarray(c.real)
    im = np.ascontiguousarray(c.imag)
    opts = _Options(_md(maxdegree), topk, max_xweight, xtrunc_period, int(bool(batched)), 0, mincoeff)
    t = _Terms()
    rc = L.ffo_pauli_propagate(_ptr(gates), len(gates), n_words, _ptr(x), _ptr(z), _ptr(re), _ptr(im),
                               len(c), ctypes.byref(opts), strategy, n_threads, ctypes.byref(t))
    if rc != 0:
        raise RuntimeError(f"oracle pauli propagate failed ({rc})")
    try:
        return _collect(t, True)
    finally:
        L.ffo_free(ctypes.byref(t))


def majorana_propagate(gates: np.ndarray, n_words: int, m: np.ndarray, c: np.ndarray,
                       maxdegree: int | None = None, mincoeff: float = 0.0) -> Result:
    L = lib()
    gates = np.ascontiguousarray(gates, dtype=MAJORANA_GATE_DTYPE)
    m = np.ascontiguousarray(m, dtype=np.uint64).reshape(-1, n_words)
    c = np.asarray(c, dtype=np.complex128).reshape(-1)
    re = np.ascontiguousarray(c.real)
    im = np.ascontiguousarray(c.imag)
    t = _Terms()
    rc = L.ffo_majorana_propagate(_ptr(gates), len(gates), n_words, _ptr(m), _ptr(re), _ptr(im), len(c),
                                  _md(maxdegree), mincoeff, ctypes.byref(t))
    if rc != 0:
        raise RuntimeError(f"oracle majorana propagate failed ({rc})")
    try:
        return _collect(t, False)
    finally:
        L.ffo_free(ctypes.byref(t))


def pauli_expect(res: Result, b_words) -> complex:
    """<b|O|b> for a qubit product state with bit mask words b_words."""
    n = len(res.c)
    W = res.x.shape[1] if n else len(b_words)
    b = np.zeros(max(W, len(b_words)), dtype=np.uint64)
    b[: len(b_words)] = b_words
    diag = ~np.any(res.x != 0, axis=1) if n else np.zeros(0, bool)
    if not n:
        return 0j
    par = np.zeros(n, dtype=np.int64)
    for w in range(W):
        v = res.z[:, w] & b[w]
        par += np.array([bin(int(q)).count("1") for q in v], dtype=np.int64)
    sign = np.where(par % 2 == 0, 1.0, -1.0)
    return complex(np.sum(res.c[diag] * sign[diag]))


def pauli_expect_native(res: Result, b_words) -> complex:
    """Same as pauli_expect, computed by the oracle library's own loop."""
    L = lib()
    n = len(res.c)
    W = res.x.shape[1] if n else 1
    t = _Terms()
    x = np.ascontiguousarray(res.x, dtype=np.uint64)
    z = np.ascontiguousarray(res.z, dtype=np.uint64)
    re = np.ascontiguousarray(res.c.real)
    im = np.ascontiguousarray(res.c.imag)
    t.n_terms, t.n_words = n, W
    t.x = x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    t.z = z.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    t.c_re = re.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t.c_im = im.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    b = np.zeros(W, dtype=np.uint64)
    b[: min(W, len(b_words))] = b_words[:W]
    out = np.zeros(2)
    L.ffo_pauli_expect(ctypes.byref(t), _ptr(b), _ptr(out))
    return complex(out[0], out[1])


def majorana_expect(res: Result, occ_words) -> complex:
    """<occ|O|occ> for a Fock product state with mode-occupation words occ_words."""
    L = lib()
    n = len(res.c)
    W = res.x.shape[1] if n else 1
    t = _Terms()
    x = np.ascontiguousarray(res.x, dtype=np.uint64)
    re = np.ascontiguousarray(res.c.real)
    im = np.ascontiguousarray(res.c.imag)
    t.n_terms, t.n_words = n, W
    t.x = x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    t.c_re = re.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t.c_im = im.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    occ = np.zeros(max(W, len(occ_words)), dtype=np.uint64)
    occ[: len(occ_words)] = occ_words
    out = np.zeros(2)
    L.ffo_majorana_expect(ctypes.byref(t), _ptr(occ), _ptr(out))
    return complex(out[0], out[1])


def max_threads() -> int:
    return int(lib().ffo_max_threads())


def rotation_counts():
    """Live terms entering each rotation of the last propagate call."""
    L = lib()
    n = L.ffo_rotation_counts(None, 0)
    out = np.zeros(n, np.int64)
    L.ffo_rotation_counts(out.ctypes.data, n)
    return out
