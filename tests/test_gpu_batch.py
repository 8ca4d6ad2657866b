"""Batched propagation (ffx_pauli_propagate_batch) in its layouts -- merged (all
instances in one term store, instance numbers in the top key bits) with one
launch per layer group step ("groups") or per window step ("windows"), and one
workgroup per instance -- against the oracle run instance by instance, and
against the device-wide engine.  The batch runs each
instance's reference schedule and truncation, so FP64 expectations agree with
the oracle to rounding of the final sum and the live term counts and work
counters are identical; instances that outgrow their slot budget (one-workgroup
layout), contain Clifford gates or ask for top-k must come back complete
through the device-wide engine."""
from __future__ import annotations

import numpy as np
import pytest

import pyffo
from fastfermion_amd import _abi, workloads as wl
from fastfermion_amd._native import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(params=["groups", "windows", "workgroup"])
def layout(request, monkeypatch):
    monkeypatch.setenv("FFX_BATCH_MERGED", "0" if request.param == "workgroup" else "1")
    if request.param == "windows":
        monkeypatch.setenv("FFX_NO_MLAYER", "1")
    return request.param


def pad(a, W):
    a = np.asarray(a, np.uint64)
    out = np.zeros((a.shape[0], W), np.uint64)
    out[:, : a.shape[1]] = a
    return out


def instances(n, seed, **kw):
    return wl.batch_kicked_ising(n, seed=seed, **kw)


def run_both(ctx, inst, W, states=None, **kw):
    items = []
    refs = []
    for i, (g, x, z, c, L, d) in enumerate(inst):
        st = None if states is None else states[i]
        items.append((g, pad(x, W), pad(z, W), c, st))
        r = pyffo.pauli_propagate(g, W, pad(x, W), pad(z, W), c,
                                  **{k: v for k, v in kw.items() if k not in ("precision", "slots")})
        b = [0] * W
        if st is not None:
            b[: len(st)] = st
        refs.append((pyffo.pauli_expect(r, b), len(r.c), r.updates))
    ex, nt, fb = ctx.pauli_propagate_batch(items, W, **kw)
    return ex, nt, fb, refs


@pytest.mark.parametrize("batched", [False, True])
def test_batch_matches_oracle_per_instance(ctx, layout, batched):
    inst = instances(24, 11, l_lo=3, l_span=4, d_lo=2, d_span=5)
    W = 1
    kw = dict(maxdegree=8, mincoeff=1e-4, batched=batched)
    # batch-end truncation lets terms pile up inside each commuting batch (the
    # RZZ(-pi/2) batches leave tiny copies alive): give those a larger budget;
    # instances that still outgrow it are exact through the fallback
    ex, nt, fb, refs = run_both(ctx, inst, W, slots=(1 << 17) if batched else 0, **kw)
    if not batched:
        assert fb == 0
    upd = 0
    for i, (e, n, u) in enumerate(refs):
        assert abs(ex[i] - e) < 1e-12, (i, ex[i], e)
        assert nt[i] == n, (i, nt[i], n)
        upd += u
    assert ctx.profile()["updates"] == upd


def test_batch_uneven_two_words_and_states(ctx, layout):
    inst = instances(16, 5, l_lo=6, l_span=5, d_lo=3, d_span=6)
    W = 2
    rng = np.random.default_rng(3)
    states = [[int(rng.integers(0, 2 ** 63)), int(rng.integers(0, 2 ** 36))] for _ in inst]
    ex, nt, fb, refs = run_both(ctx, inst, W, states=states, maxdegree=8, mincoeff=1e-4, batched=False)
    for i, (e, n, u) in enumerate(refs):
        assert abs(ex[i] - e) < 1e-12
        assert nt[i] == n


def structured_circuit(rng, nq, nblocks):
    """Blocks of rotations that the merged layout cuts into layer groups of every
    kind: rotations on disjoint random supports (some qubits left out), runs of
    Z-type or X-type strings (pairwise commuting, with repeated axes that end a
    group), and a few unstructured rotations."""
    recs = []
    for _ in range(nblocks):
        kind = int(rng.integers(0, 4))
        if kind == 0:
            perm = rng.permutation(nq).tolist()
            i = 0
            while i < nq:
                w = int(rng.integers(1, 4))
                qs = sorted(perm[i:i + w])
                i += w
                if rng.random() < 0.8:
                    recs.append(("".join(rng.choice(list("XYZ"), len(qs))), tuple(qs), float(rng.uniform(-3, 3))))
        elif kind in (1, 2):
            letter = "Z" if kind == 1 else "X"
            for _ in range(int(rng.integers(2, 14))):
                w = int(rng.integers(1, 4))
                qs = sorted(rng.choice(nq, w, replace=False).tolist())
                recs.append((letter * w, tuple(qs), float(rng.uniform(-3, 3))))
        else:
            for _ in range(int(rng.integers(1, 4))):
                w = int(rng.integers(1, 4))
                qs = sorted(rng.choice(nq, w, replace=False).tolist())
                recs.append(("".join(rng.choice(list("XYZ"), w)), tuple(qs), float(rng.uniform(-3, 3))))
    return wl._pack(recs)


@pytest.mark.parametrize("W,nq_hi", [(1, 40), (2, 120), (3, 180), (4, 240)])
@pytest.mark.parametrize("kw", [
    dict(maxdegree=8, mincoeff=1e-3, batched=False),
    dict(maxdegree=6, mincoeff=1e-4, batched=True),
    dict(maxdegree=8, mincoeff=1e-3, batched=False, max_xweight=2, xtrunc_period=2),
])
def test_batch_groups_structured_circuits(ctx, monkeypatch, W, nq_hi, kw):
    """The merged layout in layer groups on circuits with disjoint, commuting and
    unstructured runs, 1-4 key words, complex observables, both schedules and the
    X/Y cap: every instance's <O> and live term count equal the oracle's, and the
    batch's update counter equals the sum of the oracle's."""
    monkeypatch.setenv("FFX_BATCH_MERGED", "1")
    rng = np.random.default_rng(1000 * W + len(kw))
    items, refs, upd = [], [], 0
    for i in range(20):
        nq = int(rng.integers(max(4, nq_hi // 3), nq_hi + 1))
        g = structured_circuit(rng, nq, int(rng.integers(3, 12)))
        w = int(rng.integers(1, 4))
        qs = sorted(rng.choice(nq, w, replace=False).tolist())
        ops = "".join(rng.choice(list("XYZ"), w))
        coeff = 1.0 if i % 3 else 0.6 + 0.8j
        x, z, c = wl.single_pauli(ops, qs, nq, coeff)
        x, z = pad(x, W), pad(z, W)
        st = [int(v) for v in rng.integers(0, 2 ** 62, W)]
        items.append((g, x, z, c, st))
        r = pyffo.pauli_propagate(g, W, x, z, c, **kw)
        refs.append((pyffo.pauli_expect(r, st), len(r.c)))
        upd += r.updates
    ex, nt, fb = ctx.pauli_propagate_batch(items, W, **kw)
    assert fb == 0
    for i, (e, n) in enumerate(refs):
        assert abs(ex[i] - e) < 1e-12, (i, ex[i], e)
        assert nt[i] == n, (i, nt[i], n)
    assert ctx.profile()["updates"] == upd


@pytest.mark.parametrize("split", ["1", "2", "4"])
def test_batch_groups_cut_by_schedule_length(ctx, monkeypatch, split):
    """The layer-group layout cut into 1, 2 or 4 term stores by schedule length, each
    on its own stream and run side by side: every instance's <O> and live term count
    equal the oracle's whatever the cut, and the update counter equals the oracle's
    sum (the stores' counters are added)."""
    monkeypatch.setenv("FFX_BATCH_MERGED", "1")
    monkeypatch.setenv("FFX_MSPLIT", split)
    rng = np.random.default_rng(4242)
    W = 2
    items, refs, upd = [], [], 0
    for i in range(24):
        nq = int(rng.integers(20, 100))
        g = structured_circuit(rng, nq, int(rng.integers(2, 16)))  # schedules of very different lengths
        q = int(rng.integers(0, nq))
        x, z, c = wl.single_pauli("Z", [q], nq)
        x, z = pad(x, W), pad(z, W)
        items.append((g, x, z, c, None))
        r = pyffo.pauli_propagate(g, W, x, z, c, maxdegree=8, mincoeff=1e-4, batched=False)
        refs.append((pyffo.pauli_expect(r, [0, 0]), len(r.c)))
        upd += r.updates
    ex, nt, fb = ctx.pauli_propagate_batch(items, W, maxdegree=8, mincoeff=1e-4, batched=False)
    assert fb == 0
    for i, (e, n) in enumerate(refs):
        assert abs(ex[i] - e) < 1e-12, (split, i)
        assert nt[i] == n, (split, i)
    assert ctx.profile()["updates"] == upd


def test_batch_instance_numbers_that_do_not_fit(ctx, monkeypatch):
    """64 one-word instances on 60 qubits: 6 bits of instance number do not fit above
    qubit 59, so the merged layouts step aside for the one-workgroup layout, with the
    same results."""
    monkeypatch.setenv("FFX_BATCH_MERGED", "1")
    rng = np.random.default_rng(77)
    items, refs = [], []
    for i in range(64):
        nq = 60
        g = structured_circuit(rng, nq, 3)
        q = int(rng.integers(0, nq))
        x, z, c = wl.single_pauli("Z", [q], nq)
        items.append((g, x, z, c, None))
        r = pyffo.pauli_propagate(g, 1, x, z, c, maxdegree=6, mincoeff=1e-3, batched=False)
        refs.append((pyffo.pauli_expect(r, [0]), len(r.c)))
    ex, nt, fb = ctx.pauli_propagate_batch(items, 1, maxdegree=6, mincoeff=1e-3, batched=False)
    for i, (e, n) in enumerate(refs):
        assert abs(ex[i] - e) < 1e-12 and nt[i] == n, i


def test_batch_slot_budget_fallback(ctx, monkeypatch):
    """A budget far below the instances' peaks: they stop in the batch launch
    and rerun on the device-wide engine with identical results (one-workgroup
    layout; the merged store grows instead)."""
    monkeypatch.setenv("FFX_BATCH_MERGED", "0")
    inst = instances(6, 17, l_lo=7, l_span=2, d_lo=8, d_span=3)
    W = 1
    kw = dict(maxdegree=8, mincoeff=1e-5, batched=True)
    ex, nt, fb, refs = run_both(ctx, inst, W, slots=1024, **kw)
    assert fb > 0
    for i, (e, n, u) in enumerate(refs):
        assert abs(ex[i] - e) < 1e-12
        assert nt[i] == n


def test_batch_ineligible_instances(ctx, layout):
    """Clifford gates and top-k go to the device-wide engine; mixed with normal
    instances in one call."""
    inst = instances(4, 23, l_lo=3, l_span=2, d_lo=2, d_span=3)
    W = 1
    items, refs = [], []
    for j, (g, x, z, c, L, d) in enumerate(inst):
        if j == 1:
            g = np.concatenate([g[:5], _abi.clifford_record(_abi.H, 0, 0), g[5:]])
        items.append((g, x, z, c, None))
        r = pyffo.pauli_propagate(g, W, x, z, c, maxdegree=8, mincoeff=1e-4)
        refs.append(pyffo.pauli_expect(r, [0]))
    ex, nt, fb = ctx.pauli_propagate_batch(items, W, maxdegree=8, mincoeff=1e-4)
    assert fb >= 1  # the Clifford instance, plus any that outgrew the default budget
    for i, e in enumerate(refs):
        assert abs(ex[i] - e) < 1e-12
    ex2, _, fb2 = ctx.pauli_propagate_batch(items, W, maxdegree=8, mincoeff=1e-4, topk=50)
    assert fb2 == len(items)


@pytest.mark.parametrize("prec,tol", [(_abi.FP32C, 1e-9), (_abi.FP32, 1e-4)])
def test_batch_precisions(ctx, layout, prec, tol):
    inst = instances(12, 29, l_lo=4, l_span=4, d_lo=3, d_span=6)
    W = 1
    items = [(g, x, z, c, None) for (g, x, z, c, L, d) in inst]
    kw = dict(maxdegree=8, mincoeff=1e-4, batched=False)
    e64, _, _ = ctx.pauli_propagate_batch(items, W, **kw)
    e32, _, _ = ctx.pauli_propagate_batch(items, W, precision=prec, **kw)
    assert np.max(np.abs(e64 - e32)) < tol


def test_batch_equals_single_engine(ctx, layout):
    inst = instances(5, 31, l_lo=5, l_span=3, d_lo=4, d_span=4)
    W = 1
    items = [(g, x, z, c, None) for (g, x, z, c, L, d) in inst]
    kw = dict(maxdegree=8, mincoeff=1e-4, batched=True)
    ex, nt, _ = ctx.pauli_propagate_batch(items, W, **kw)
    for i, (g, x, z, c, _) in enumerate(items):
        res, e = ctx.pauli_propagate(g, W, x, z, c, state=[0], **kw)
        assert abs(ex[i] - e) < 1e-12
        assert nt[i] == len(res)


def test_batch_empty_and_trivial(ctx, layout):
    ex, nt, fb = ctx.pauli_propagate_batch([], 1)
    assert len(ex) == 0 and fb == 0
    x, z, c = wl.single_pauli("Z", [0], 4)
    ex, nt, fb = ctx.pauli_propagate_batch([(np.zeros(0, _abi.PAULI_GATE_DTYPE), x, z, c, None)], 1)
    assert ex[0] == 1.0 and nt[0] == 1


def test_graph_replays_match_direct_launches(ctx, monkeypatch):
    """A batch whose pool layout repeats (same instances, new angles: an angle
    sweep) replays the captured HIP graph.  Every replay, with eager work and
    synchronisations between them, must equal the direct launch bit for bit
    (one-workgroup layout, the one that is captured)."""
    import math

    monkeypatch.setenv("FFX_BATCH_MERGED", "0")
    W = 1
    base = instances(32, 21, l_lo=3, l_span=4, d_lo=2, d_span=6)
    kw = dict(maxdegree=8, mincoeff=1e-4, batched=False)

    def sweep(scale):
        items = []
        for (g, x, z, c, L, d) in base:
            g2 = g.copy()
            rx = (g2["x"][:, 0] != 0) & (g2["z"][:, 0] == 0)
            g2["theta"][rx] = g2["theta"][rx] * scale
            items.append((g2, pad(x, W), pad(z, W), c, None))
        return items

    monkeypatch.setenv("FFX_BATCH_NO_GRAPH", "1")
    direct = [ctx.pauli_propagate_batch(sweep(s), W, **kw) for s in (1.0, 0.7, 1.3)]
    monkeypatch.delenv("FFX_BATCH_NO_GRAPH")
    r0 = ctx.batch_graph_replays()
    for rep in range(3):
        for s, want in zip((1.0, 0.7, 1.3), direct):
            got = ctx.pauli_propagate_batch(sweep(s), W, **kw)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]) and got[2] == want[2]
            # eager work on the same context between replays
            g, x, z, c, L, d = base[rep]
            ctx.pauli_propagate(g, W, pad(x, W), pad(z, W), c, download=False, want_expect=True, **kw)
    assert ctx.batch_graph_replays() - r0 >= 8  # the first call of the run captures, the rest replay
    assert all(math.isfinite(v.real) for v in direct[0][0])


def test_autotune_keeps_results(ctx):
    """The autotuned launch (layout, compaction interval, group size, workgroup
    size, window length, slot budget) changes only speed: FP64 term counts stay
    identical to the default launch and expectations agree to rounding of their
    sums."""
    W = 1
    inst = instances(40, 31, l_lo=3, l_span=5, d_lo=3, d_span=8)
    items = [(g, pad(x, W), pad(z, W), c, None) for (g, x, z, c, L, d) in inst]
    kw = dict(maxdegree=8, mincoeff=1e-4, batched=False)
    want = ctx.pauli_propagate_batch(items, W, **kw)
    t = ctx.batch_autotune(items, W, sample=16, **kw)
    assert t["block"] in (256, 512) and 1 <= t["window"] <= 4 and t["slots"] >= 1024
    assert t["merged"] in (0, 1, 2)
    if t["merged"] == 1:
        assert t["compact_every"] in (2, 4) and t["group_max"] in (6, 8) and t["split"] in (1, 2, 4)
    assert t["candidates"] == 21 and t["sample"] == 16 and t["ms"] > 0
    got = ctx.pauli_propagate_batch(items, W, **kw)
    assert np.max(np.abs(got[0] - want[0])) < 1e-13 and np.array_equal(got[1], want[1])
