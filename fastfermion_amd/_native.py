"""ctypes binding of the C ABI in libffamd.so (include/fastfermion_amd.h).

This is the array-level entry to the device engine used by bench.py and the
parity tests; the object-level drop-in API (PauliString, ROT, propagate, ...)
lives in the compiled module fastfermion_amd._ffcore, which calls the same ABI.
There is no CPU fallback: every call raises when no HIP device is usable.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from ._abi import FP64, MAJORANA_GATE_DTYPE, PAULI_GATE_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffamd.so")

STATUS = {0: "ok", -1: "invalid argument", -2: "out of memory", -3: "HIP error", -4: "no HIP device",
          -5: "capacity exceeded"}

EXPORTED = ("ffx_ctx_create", "ffx_ctx_destroy", "ffx_last_error", "ffx_pauli_propagate",
            "ffx_majorana_propagate", "ffx_expect_last", "ffx_terms_free", "ffx_profile_get",
            "ffx_device_count", "ffx_set_kernel_timing", "ffx_pauli_propagate_batch",
            "ffx_pauli_propagate_pruned", "ffx_majorana_propagate_pruned", "ffx_pauli_contrib_log",
            "ffx_majorana_contrib_log", "ffx_pruner_init", "ffx_pruner_set_params", "ffx_pruner_get_params",
            "ffx_pruner_forward", "ffx_pruner_loss_grad", "ffx_pruner_train", "ffx_batch_autotune",
            "ffx_batch_graph_replays", "ffx_pauli_propagate_noalloc")

# maxdegree value that disables the partner degree filter (FFX_NO_DEGREE_LIMIT)
NO_DEGREE_LIMIT = 0x7FFFFFFF


def _md(maxdegree):
    """None: no filter; an integer is compared literally, as the reference does
    (src/pauli/propagate.h:82, src/majorana_gates.h:94): a negative one drops every partner."""
    return NO_DEGREE_LIMIT if maxdegree is None else int(maxdegree)


PRUNER_FEATURES = 64
PRUNER_NPARAMS = 24961


class Options(ctypes.Structure):
    _fields_ = [("maxdegree", ctypes.c_int32), ("topk", ctypes.c_int32),
                ("max_xweight", ctypes.c_int32), ("xtrunc_period", ctypes.c_int32),
                ("batched", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("mincoeff", ctypes.c_double)]


class PruneOpts(ctypes.Structure):
    """ffx_prune_opts: pruned-mode switches and the site geometry of the term features."""
    _fields_ = [("clifford_tol", ctypes.c_double), ("log10_tau", ctypes.c_float),
                ("prune_period", ctypes.c_int32), ("rot_per_layer", ctypes.c_int32),
                ("obs_site", ctypes.c_int32), ("row", ctypes.c_uint8 * 256), ("col", ctypes.c_uint8 * 256)]


def prune_opts(row, col, obs_site=0, clifford_tol=0.0, prune_period=0, log10_tau=-30.0, rot_per_layer=1):
    """PruneOpts from per-index lattice rows/columns (qubits, or Majorana indices)."""
    po = PruneOpts()
    po.clifford_tol = float(clifford_tol)
    po.log10_tau = float(log10_tau)
    po.prune_period = int(prune_period)
    po.rot_per_layer = int(rot_per_layer)
    po.obs_site = int(obs_site)
    for i, (r, c) in enumerate(zip(row, col)):
        po.row[i] = int(r)
        po.col[i] = int(c)
    return po


class Terms(ctypes.Structure):
    _fields_ = [("n_terms", ctypes.c_int64), ("n_words", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("x", ctypes.POINTER(ctypes.c_uint64)), ("z", ctypes.POINTER(ctypes.c_uint64)),
                ("c_re", ctypes.POINTER(ctypes.c_double)), ("c_im", ctypes.POINTER(ctypes.c_double))]


class Instance(ctypes.Structure):
    _fields_ = [("gates", ctypes.c_void_p), ("n_gates", ctypes.c_int64), ("x", ctypes.c_void_p),
                ("z", ctypes.c_void_p), ("c_re", ctypes.c_void_p), ("c_im", ctypes.c_void_p),
                ("n_terms", ctypes.c_int64), ("basis_state", ctypes.c_uint64 * 4)]


# numpy image of Instance (same field offsets), for filling many records at once
_INSTANCE_DTYPE = np.dtype([("gates", "<u8"), ("n_gates", "<i8"), ("x", "<u8"), ("z", "<u8"), ("c_re", "<u8"),
                            ("c_im", "<u8"), ("n_terms", "<i8"), ("basis_state", "<u8", (4,))])
assert _INSTANCE_DTYPE.itemsize == ctypes.sizeof(Instance)
for _f, _t in Instance._fields_:
    assert _INSTANCE_DTYPE.fields[_f][1] == getattr(Instance, _f).offset


class BatchTuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("window", ctypes.c_int32), ("slots", ctypes.c_int64),
                ("ms", ctypes.c_double), ("candidates", ctypes.c_int32), ("sample", ctypes.c_int32),
                ("merged", ctypes.c_int32), ("compact_every", ctypes.c_int32), ("group_max", ctypes.c_int32),
                ("split", ctypes.c_int32)]


class Profile(ctypes.Structure):
    _fields_ = [("n_gates", ctypes.c_int64), ("updates", ctypes.c_int64), ("peak_terms", ctypes.c_int64),
                ("launches", ctypes.c_int64), ("seconds_total", ctypes.c_double),
                ("seconds_upload", ctypes.c_double), ("seconds_propagate", ctypes.c_double),
                ("seconds_download", ctypes.c_double), ("outs", ctypes.c_int64),
                ("compactions", ctypes.c_int64), ("run_launches", ctypes.c_int64),
                ("seconds_kernel", ctypes.c_double), ("timed_launches", ctypes.c_int64),
                ("seconds_enqueue", ctypes.c_double)]


_lib = None
_lib_lock = threading.Lock()


def load():
    """Load libffamd.so (raises OSError if it was not built)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise OSError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
            L = ctypes.CDLL(LIB_PATH)
            vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
            L.ffx_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
            L.ffx_ctx_destroy.argtypes = [vp]
            L.ffx_last_error.argtypes = [vp]
            L.ffx_last_error.restype = ctypes.c_char_p
            L.ffx_pauli_propagate.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp, i64, ctypes.POINTER(Options),
                                              ctypes.POINTER(Terms), vp, vp]
            L.ffx_majorana_propagate.argtypes = [vp, vp, i64, i32, vp, vp, vp, i64, i32, dbl, i32,
                                                 ctypes.POINTER(Terms), vp, vp]
            L.ffx_expect_last.argtypes = [vp, vp, vp]
            L.ffx_terms_free.argtypes = [ctypes.POINTER(Terms)]
            L.ffx_profile_get.argtypes = [vp, ctypes.POINTER(Profile)]
            L.ffx_device_count.argtypes = []
            L.ffx_set_kernel_timing.argtypes = [vp, ctypes.c_int]
            L.ffx_pauli_propagate_batch.argtypes = [vp, ctypes.POINTER(Instance), i32, i32, ctypes.POINTER(Options),
                                                    i64, vp, vp, vp]
            PO = ctypes.POINTER(PruneOpts)
            L.ffx_pauli_propagate_pruned.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp, i64, ctypes.POINTER(Options),
                                                     PO, ctypes.POINTER(Terms), vp, vp]
            L.ffx_majorana_propagate_pruned.argtypes = [vp, vp, i64, i32, vp, vp, vp, i64, i32, dbl, i32, PO,
                                                        ctypes.POINTER(Terms), vp, vp]
            L.ffx_pauli_contrib_log.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp, i64, ctypes.POINTER(Options), PO, vp,
                                                i32, i64, vp, vp, vp, vp, i32, vp]
            L.ffx_majorana_contrib_log.argtypes = [vp, vp, i64, i32, vp, vp, vp, i64, i32, dbl, PO, vp, i32, i64, vp,
                                                   vp, vp, vp, i32, vp]
            L.ffx_batch_autotune.argtypes = [vp, ctypes.POINTER(Instance), i32, i32, ctypes.POINTER(Options), i32,
                                             ctypes.POINTER(BatchTuning)]
            L.ffx_batch_graph_replays.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
            L.ffx_pauli_propagate_noalloc.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp, i64, ctypes.POINTER(Options),
                                                      vp, vp, vp, vp, vp, vp, i64, ctypes.POINTER(ctypes.c_int64)]
            L.ffx_pruner_init.argtypes = [vp, ctypes.c_uint64]
            L.ffx_pruner_set_params.argtypes = [vp, vp, i64]
            L.ffx_pruner_get_params.argtypes = [vp, vp, i64]
            L.ffx_pruner_forward.argtypes = [vp, vp, i64, vp]
            L.ffx_pruner_loss_grad.argtypes = [vp, vp, vp, i64, ctypes.c_float, vp, vp]
            L.ffx_pruner_train.argtypes = [vp, vp, vp, i64, i32, i32, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                                           vp]
            _lib = L
    return _lib


def device_count() -> int:
    return int(load().ffx_device_count())


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Result:
    """Propagated observable: term-major uint64 keys and complex coefficients."""

    def __init__(self, x, z, c):
        self.x, self.z, self.c = x, z, c

    def __len__(self):
        return len(self.c)


class Context:
    """One HIP device context (stream + device memory pools) of libffamd.so."""

    def __init__(self, device: int = 0):
        self._L = load()
        h = ctypes.c_void_p()
        rc = self._L.ffx_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"fastfermion_amd: cannot create a HIP context on device {device} "
                               f"({STATUS.get(rc, rc)})")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.ffx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            msg = self._L.ffx_last_error(self._h).decode(errors="replace")
            if rc == -2:
                raise MemoryError(msg)
            raise RuntimeError(f"fastfermion_amd: {STATUS.get(rc, rc)}: {msg}")

    @staticmethod
    def _collect(t: Terms, pauli: bool) -> Result:
        n, W = t.n_terms, t.n_words
        if n == 0:
            e = np.zeros((0, W), np.uint64)
            return Result(e, e.copy() if pauli else None, np.zeros(0, np.complex128))
        x = np.ctypeslib.as_array(t.x, shape=(n * W,)).reshape(n, W).copy()
        z = np.ctypeslib.as_array(t.z, shape=(n * W,)).reshape(n, W).copy() if pauli else None
        re = np.ctypeslib.as_array(t.c_re, shape=(n,)).copy()
        im = np.ctypeslib.as_array(t.c_im, shape=(n,)).copy()
        return Result(x, z, re + 1j * im)

    def pauli_propagate(self, gates, n_words, x, z, c, maxdegree=None, mincoeff=0.0, topk=0, max_xweight=-1,
                        xtrunc_period=1, batched=True, precision=FP64, download=True, state=None,
                        want_expect=False):
        gates = np.ascontiguousarray(gates, dtype=PAULI_GATE_DTYPE)
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, n_words)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, n_words)
        c = np.asarray(c, dtype=np.complex128).reshape(-1)
        re = np.ascontiguousarray(c.real)
        im = np.ascontiguousarray(c.imag) if np.any(c.imag != 0) else None
        opts = Options(_md(maxdegree), topk, max_xweight, xtrunc_period, int(bool(batched)), precision, mincoeff)
        t = Terms()
        ex = np.zeros(2)
        st = None
        if state is not None:
            st = np.zeros(4, np.uint64)
            st[: len(state)] = state
        rc = self._L.ffx_pauli_propagate(self._h, _p(gates), len(gates), n_words, _p(x), _p(z), _p(re), _p(im),
                                         len(c), ctypes.byref(opts), ctypes.byref(t) if download else None,
                                         _p(st), _p(ex) if (want_expect or state is not None) else None)
        self._check(rc)
        res = None
        if download:
            try:
                res = self._collect(t, True)
            finally:
                self._L.ffx_terms_free(ctypes.byref(t))
        if want_expect or state is not None:
            return res, complex(ex[0], ex[1])
        return res

    def pauli_propagate_noalloc(self, gates, n_words, x, z, c, maxdegree=None, mincoeff=0.0, topk=0, max_xweight=-1,
                                xtrunc_period=1, batched=True, precision=FP64, state=None, out_cap=None):
        """ffx_pauli_propagate_noalloc: no device allocation; the terms come back in
        arrays of out_cap terms allocated here (None: 1 << 16; 0: expectation and
        term count only).  Returns (Result or None, <b|O|b>, live term count)."""
        gates = np.ascontiguousarray(gates, dtype=PAULI_GATE_DTYPE)
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, n_words)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, n_words)
        c = np.asarray(c, dtype=np.complex128).reshape(-1)
        re = np.ascontiguousarray(c.real)
        im = np.ascontiguousarray(c.imag) if np.any(c.imag != 0) else None
        opts = Options(_md(maxdegree), topk, max_xweight, xtrunc_period, int(bool(batched)), precision, mincoeff)
        cap = 1 << 16 if out_cap is None else int(out_cap)
        ox = np.zeros((cap, n_words), np.uint64)
        oz = np.zeros((cap, n_words), np.uint64)
        ore = np.zeros(cap)
        oim = np.zeros(cap)
        ex = np.zeros(2)
        st = np.zeros(4, np.uint64)
        if state is not None:
            st[: len(state)] = state
        n = ctypes.c_int64(0)
        rc = self._L.ffx_pauli_propagate_noalloc(self._h, _p(gates), len(gates), n_words, _p(x), _p(z), _p(re), _p(im),
                                                 len(c), ctypes.byref(opts), _p(st), _p(ex),
                                                 _p(ox) if cap else None, _p(oz) if cap else None,
                                                 _p(ore) if cap else None, _p(oim) if cap else None, cap,
                                                 ctypes.byref(n))
        self.last_noalloc_terms = int(n.value)
        self._check(rc)
        k = int(n.value)
        res = Result(ox[:k].copy(), oz[:k].copy(), ore[:k] + 1j * oim[:k]) if cap else None
        return res, complex(ex[0], ex[1]), k

    def pauli_propagate_batch(self, instances, n_words, maxdegree=None, mincoeff=0.0, topk=0, max_xweight=-1,
                              xtrunc_period=1, batched=True, precision=FP64, slots=0):
        """instances: iterable of (gates, x, z, c, state) with state a list of
        up to 4 uint64 words (or None for |0...0>).  Returns (expectations
        complex[n], live term counts int64[n], number of instances that ran on
        the device-wide engine)."""
        arr, keep = self._instances(instances, n_words)
        n = len(instances)
        opts = Options(_md(maxdegree), topk, max_xweight, xtrunc_period, int(bool(batched)), precision, mincoeff)
        ex = np.zeros(2 * max(n, 1))
        nt = np.zeros(max(n, 1), np.int64)
        fb = ctypes.c_int32(0)
        self._check(self._L.ffx_pauli_propagate_batch(self._h, arr, n, n_words, ctypes.byref(opts), slots, _p(ex),
                                                      _p(nt), ctypes.byref(fb)))
        del keep
        return ex[0:2 * n:2] + 1j * ex[1:2 * n:2], nt[:n], int(fb.value)

    def batch_autotune(self, instances, n_words, sample=0, maxdegree=None, mincoeff=0.0, topk=0, max_xweight=-1,
                       xtrunc_period=1, batched=True, precision=FP64):
        """Autotune the batched launch on the whole batch (sample = 0) or on the
        `sample` longest instances; later batches with slots=0 use the result.
        Returns the chosen configuration as a dict."""
        arr, keep = self._instances(instances, n_words)
        opts = Options(_md(maxdegree), topk, max_xweight, xtrunc_period, int(bool(batched)), precision, mincoeff)
        t = BatchTuning()
        self._check(self._L.ffx_batch_autotune(self._h, arr, len(instances), n_words, ctypes.byref(opts), sample,
                                               ctypes.byref(t)))
        del keep
        return {k: getattr(t, k) for k, _ in BatchTuning._fields_}

    def batch_graph_replays(self) -> int:
        v = ctypes.c_int64(0)
        self._check(self._L.ffx_batch_graph_replays(self._h, ctypes.byref(v)))
        return int(v.value)

    @staticmethod
    def _instances(instances, n_words):
        """ffx_instance records of a batch, filled column-wise in a numpy array of the
        same layout (a ctypes field assignment per record costs ~10 us); the
        observables of all instances are packed into one key / coefficient array each,
        the gate tables stay where they are.  Returns the record pointer and the
        arrays it points into, which must outlive the call."""
        n = len(instances)
        rec = np.zeros(max(n, 1), dtype=_INSTANCE_DTYPE)
        if n == 0:
            return rec.ctypes.data_as(ctypes.POINTER(Instance)), [rec]
        gl, xl, zl, cl = [], [], [], []
        gptr = np.zeros(n, np.uint64)
        ng = np.zeros(n, np.int64)
        for i, (gates, x, z, c, state) in enumerate(instances):
            gates = np.ascontiguousarray(gates, dtype=PAULI_GATE_DTYPE)
            gl.append(gates)
            gptr[i] = gates.ctypes.data
            ng[i] = len(gates)
            xl.append(np.asarray(x, dtype=np.uint64).reshape(-1, n_words))
            zl.append(np.asarray(z, dtype=np.uint64).reshape(-1, n_words))
            cl.append(np.asarray(c, dtype=np.complex128).reshape(-1))
            if state:
                for w, v in enumerate(list(state)[:4]):
                    rec["basis_state"][i, w] = int(v)
        X = np.ascontiguousarray(np.concatenate(xl))
        Z = np.ascontiguousarray(np.concatenate(zl))
        C = np.concatenate(cl)
        RE = np.ascontiguousarray(C.real)
        IM = np.ascontiguousarray(C.imag)
        nt = np.array([len(c) for c in cl], np.int64)
        toff = np.concatenate([[0], np.cumsum(nt)[:-1]]).astype(np.uint64)
        nz = np.concatenate([[0], np.cumsum(IM != 0)])
        has_im = (nz[toff.astype(np.int64) + nt] - nz[toff.astype(np.int64)]) > 0
        rec["gates"][:n] = gptr
        rec["n_gates"][:n] = ng
        rec["x"][:n] = np.uint64(X.ctypes.data) + toff * np.uint64(8 * n_words)
        rec["z"][:n] = np.uint64(Z.ctypes.data) + toff * np.uint64(8 * n_words)
        rec["c_re"][:n] = np.uint64(RE.ctypes.data) + toff * np.uint64(8)
        rec["c_im"][:n] = np.where(has_im, np.uint64(IM.ctypes.data) + toff * np.uint64(8), np.uint64(0))
        rec["n_terms"][:n] = nt
        return rec.ctypes.data_as(ctypes.POINTER(Instance)), [rec, X, Z, RE, IM] + gl

    def majorana_propagate(self, gates, n_words, m, c, maxdegree=None, mincoeff=0.0, precision=FP64,
                           download=True, occ=None, want_expect=False):
        gates = np.ascontiguousarray(gates, dtype=MAJORANA_GATE_DTYPE)
        m = np.ascontiguousarray(m, dtype=np.uint64).reshape(-1, n_words)
        c = np.asarray(c, dtype=np.complex128).reshape(-1)
        re = np.ascontiguousarray(c.real)
        im = np.ascontiguousarray(c.imag) if np.any(c.imag != 0) else None
        t = Terms()
        ex = np.zeros(2)
        oc = None
        if occ is not None:
            oc = np.zeros(4, np.uint64)
            oc[: len(occ)] = occ
        rc = self._L.ffx_majorana_propagate(self._h, _p(gates), len(gates), n_words, _p(m), _p(re), _p(im),
                                            len(c), _md(maxdegree), mincoeff, precision,
                                            ctypes.byref(t) if download else None, _p(oc),
                                            _p(ex) if (want_expect or occ is not None) else None)
        self._check(rc)
        res = None
        if download:
            try:
                res = self._collect(t, False)
            finally:
                self._L.ffx_terms_free(ctypes.byref(t))
        if want_expect or occ is not None:
            return res, complex(ex[0], ex[1])
        return res

    # ---- pruned mode, contribution logs, learned pruner ----
    @staticmethod
    def _opts(maxdegree, topk, max_xweight, xtrunc_period, batched, precision, mincoeff):
        return Options(_md(maxdegree), topk, max_xweight, xtrunc_period, int(bool(batched)), precision, mincoeff)

    def pauli_propagate_pruned(self, gates, n_words, x, z, c, po, maxdegree=None, mincoeff=0.0, topk=0, max_xweight=-1,
                               xtrunc_period=1, batched=True, precision=FP64, download=True, state=None,
                               want_expect=False):
        """pauli_propagate with the pruned mode of `po` (a PruneOpts)."""
        gates = np.ascontiguousarray(gates, dtype=PAULI_GATE_DTYPE)
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, n_words)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, n_words)
        c = np.asarray(c, dtype=np.complex128).reshape(-1)
        re = np.ascontiguousarray(c.real)
        im = np.ascontiguousarray(c.imag) if np.any(c.imag != 0) else None
        opts = self._opts(maxdegree, topk, max_xweight, xtrunc_period, batched, precision, mincoeff)
        t = Terms()
        ex = np.zeros(2)
        st = None
        if state is not None:
            st = np.zeros(4, np.uint64)
            st[: len(state)] = state
        rc = self._L.ffx_pauli_propagate_pruned(self._h, _p(gates), len(gates), n_words, _p(x), _p(z), _p(re),
                                                _p(im), len(c), ctypes.byref(opts),
                                                ctypes.byref(po) if po is not None else None,
                                                ctypes.byref(t) if download else None, _p(st),
                                                _p(ex) if (want_expect or state is not None) else None)
        self._check(rc)
        res = None
        if download:
            try:
                res = self._collect(t, True)
            finally:
                self._L.ffx_terms_free(ctypes.byref(t))
        if want_expect or state is not None:
            return res, complex(ex[0], ex[1])
        return res

    def majorana_propagate_pruned(self, gates, n_words, m, c, po, maxdegree=None, mincoeff=0.0, precision=FP64,
                                  download=True, occ=None, want_expect=False):
        gates = np.ascontiguousarray(gates, dtype=MAJORANA_GATE_DTYPE)
        m = np.ascontiguousarray(m, dtype=np.uint64).reshape(-1, n_words)
        c = np.asarray(c, dtype=np.complex128).reshape(-1)
        re = np.ascontiguousarray(c.real)
        im = np.ascontiguousarray(c.imag) if np.any(c.imag != 0) else None
        t = Terms()
        ex = np.zeros(2)
        oc = None
        if occ is not None:
            oc = np.zeros(4, np.uint64)
            oc[: len(occ)] = occ
        rc = self._L.ffx_majorana_propagate_pruned(self._h, _p(gates), len(gates), n_words, _p(m), _p(re), _p(im),
                                                   len(c), _md(maxdegree), mincoeff, precision,
                                                   ctypes.byref(po) if po is not None else None,
                                                   ctypes.byref(t) if download else None, _p(oc),
                                                   _p(ex) if (want_expect or occ is not None) else None)
        self._check(rc)
        res = None
        if download:
            try:
                res = self._collect(t, False)
            finally:
                self._L.ffx_terms_free(ctypes.byref(t))
        if want_expect or occ is not None:
            return res, complex(ex[0], ex[1])
        return res

    def _log_out(self, max_rows, n_points):
        feats = np.zeros((max(max_rows, 1), PRUNER_FEATURES), np.float32)
        targ = np.zeros(max(max_rows, 1), np.float32)
        sums = np.zeros(2 * n_points)
        return feats, targ, sums

    def pauli_contrib_log(self, gates, n_words, x, z, c, po, log_every=1, max_rows=1 << 20, maxdegree=None,
                          mincoeff=0.0, max_xweight=-1, xtrunc_period=1, batched=True, state=None):
        """Term-contribution log of one plain FP64 propagation: (features float32[n, 64],
        targets float32[n] = clamp(log10(|v| + 1e-12), -12, 0)/4, sums complex[points]
        = sum_q c_q v_q at every logged step)."""
        gates = np.ascontiguousarray(gates, dtype=PAULI_GATE_DTYPE)
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, n_words)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, n_words)
        c = np.asarray(c, dtype=np.complex128).reshape(-1)
        re = np.ascontiguousarray(c.real)
        im = np.ascontiguousarray(c.imag) if np.any(c.imag != 0) else None
        opts = self._opts(maxdegree, 0, max_xweight, xtrunc_period, batched, FP64, mincoeff)
        npts = len(gates) // max(log_every, 1) + 2
        feats, targ, sums = self._log_out(max_rows, npts)
        st = np.zeros(4, np.uint64)
        if state is not None:
            st[: len(state)] = state
        nr, npo = ctypes.c_int64(0), ctypes.c_int32(0)
        self._check(self._L.ffx_pauli_contrib_log(self._h, _p(gates), len(gates), n_words, _p(x), _p(z), _p(re),
                                                  _p(im), len(c), ctypes.byref(opts), ctypes.byref(po), _p(st),
                                                  log_every, max_rows, _p(feats), _p(targ), ctypes.byref(nr),
                                                  _p(sums), npts, ctypes.byref(npo)))
        n, k = nr.value, npo.value
        return feats[:n], targ[:n], sums[0:2 * k:2] + 1j * sums[1:2 * k:2]

    def majorana_contrib_log(self, gates, n_words, m, c, po, log_every=1, max_rows=1 << 20, maxdegree=None,
                             mincoeff=0.0, occ=None):
        gates = np.ascontiguousarray(gates, dtype=MAJORANA_GATE_DTYPE)
        m = np.ascontiguousarray(m, dtype=np.uint64).reshape(-1, n_words)
        c = np.asarray(c, dtype=np.complex128).reshape(-1)
        re = np.ascontiguousarray(c.real)
        im = np.ascontiguousarray(c.imag) if np.any(c.imag != 0) else None
        npts = len(gates) // max(log_every, 1) + 2
        feats, targ, sums = self._log_out(max_rows, npts)
        oc = np.zeros(4, np.uint64)
        if occ is not None:
            oc[: len(occ)] = occ
        nr, npo = ctypes.c_int64(0), ctypes.c_int32(0)
        self._check(self._L.ffx_majorana_contrib_log(self._h, _p(gates), len(gates), n_words, _p(m), _p(re), _p(im),
                                                     len(c), _md(maxdegree), mincoeff, ctypes.byref(po), _p(oc), log_every,
                                                     max_rows, _p(feats), _p(targ), ctypes.byref(nr), _p(sums), npts,
                                                     ctypes.byref(npo)))
        n, k = nr.value, npo.value
        return feats[:n], targ[:n], sums[0:2 * k:2] + 1j * sums[1:2 * k:2]

    def pruner_init(self, seed: int = 1) -> None:
        self._check(self._L.ffx_pruner_init(self._h, ctypes.c_uint64(seed)))

    def pruner_set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        self._check(self._L.ffx_pruner_set_params(self._h, _p(p), len(p)))

    def pruner_get_params(self):
        p = np.zeros(PRUNER_NPARAMS, np.float32)
        self._check(self._L.ffx_pruner_get_params(self._h, _p(p), len(p)))
        return p

    def pruner_forward(self, feats):
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, PRUNER_FEATURES)
        out = np.zeros(len(f), np.float32)
        self._check(self._L.ffx_pruner_forward(self._h, _p(f), len(f), _p(out)))
        return out

    def pruner_loss_grad(self, feats, targets, under_weight=1.0):
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, PRUNER_FEATURES)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        g = np.zeros(PRUNER_NPARAMS, np.float32)
        loss = ctypes.c_double(0.0)
        self._check(self._L.ffx_pruner_loss_grad(self._h, _p(f), _p(t), len(f), ctypes.c_float(under_weight),
                                                 ctypes.byref(loss), _p(g)))
        return loss.value, g

    def pruner_train(self, feats, targets, epochs=4, batch=4096, lr=1e-3, seed=7, under_weight=1.0):
        """Adam on the device; returns the loss of every epoch."""
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, PRUNER_FEATURES)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        hist = np.zeros(max(epochs, 1))
        self._check(self._L.ffx_pruner_train(self._h, _p(f), _p(t), len(f), epochs, batch, ctypes.c_float(lr),
                                             ctypes.c_float(under_weight), ctypes.c_uint64(seed), _p(hist)))
        return hist[:epochs]

    def expect_last(self, state=None) -> complex:
        st = np.zeros(4, np.uint64)
        if state is not None:
            st[: len(state)] = state
        ex = np.zeros(2)
        self._check(self._L.ffx_expect_last(self._h, _p(st), _p(ex)))
        return complex(ex[0], ex[1])

    def set_kernel_timing(self, every: int) -> None:
        """Bracket every `every`-th propagation kernel launch with HIP events
        (0: off); profile() then reports seconds_kernel over timed_launches."""
        self._check(self._L.ffx_set_kernel_timing(self._h, int(every)))

    def profile(self) -> dict:
        p = Profile()
        self._check(self._L.ffx_profile_get(self._h, ctypes.byref(p)))
        return {k: getattr(p, k) for k, _ in Profile._fields_}


_default = None


def default_context() -> Context:
    """Process-wide context on device 0 (created on first use)."""
    global _default
    if _default is None:
        _default = Context(0)
    return _default
