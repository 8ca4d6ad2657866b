"""Parity at the workloads' full sizes (device vs the CPU oracle).

* 10x10 kicked Ising, 20 layers, Z_55, maxdegree 8, mincoeff 1e-5, FP64, both
  schedules: keys identical and coefficients bit-identical to the oracle's
  key-sharded run (src/pauli/propagate.h:157-264, 297-493).
* The same circuit in FP32 and FP32 compensated: |d<O>|, relative coefficient
  error on common keys and the key-set symmetric difference against FP64.
* 512 uneven kicked-Ising instances (L 6..12, so up to 144 qubits and 3 key
  words; depth 4..24) batched in every layout (one term store for all in
  layer-group or window steps, one workgroup each), FP64 and FP32 compensated:
  every instance's <O> and live term count against the oracle run per instance.
* The 6x6 Fermi-Hubbard Majorana circuit (20 Trotter steps, 6960 rotations) in
  both term orders, bit-identical to the serial Majorana oracle
  (src/majorana_propagation.h:18-33); and the learned pruner on the 30-layer
  kicked-Ising and the Fermi-Hubbard circuits within |d<O>| <= 1e-3 of the oracle.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import pyffo
from _helpers import assert_same_terms, canon
from fastfermion_amd import _abi, pruner, workloads as wl
from fastfermion_amd._native import Context, prune_opts

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
KI_OPTS = dict(maxdegree=8, mincoeff=1e-5)  # the 10x10 kicked-Ising truncation


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c2():
    gates, _ = wl.kicked_ising(10, 20)
    x, z, c = wl.single_pauli("Z", [55], 100)
    return gates, x, z, c


@pytest.fixture(scope="module")
def c2_oracle(c2):
    gates, x, z, c = c2
    return {b: pyffo.pauli_propagate(gates, 2, x, z, c, batched=b, strategy=pyffo.SHARDED, n_threads=THREADS, **KI_OPTS)
            for b in (True, False)}


@pytest.mark.parametrize("batched", [True, False])
def test_c2_full_size_bit_exact(ctx, c2, c2_oracle, batched):
    gates, x, z, c = c2
    ref = c2_oracle[batched]
    dev = ctx.pauli_propagate(gates, 2, x, z, c, batched=batched, **KI_OPTS)
    p = ctx.profile()
    print(f"kicked Ising 10x10x20 batched={batched}: {len(dev)} terms, {p['updates']} updates (oracle {ref.updates}), "
          f"peak {p['peak_terms']}", flush=True)
    assert_same_terms(dev, ref, exact=True, what=f"kicked Ising 10x10x20 batched={batched}")
    assert p["updates"] == ref.updates
    assert abs(ctx.expect_last([0, 0]) - pyffo.pauli_expect(ref, [0, 0])) < 1e-15


def precision_errors(dev, ref):
    """|d<O>| on |0...0>, max and median relative error of the coefficients on the
    common keys, keys on one side only (and their largest |c|)."""
    ka, ca = canon(dev.x, dev.z, dev.c)
    kb, cb = canon(ref.x, ref.z, ref.c)
    da = {tuple(map(int, ka[i])): ca[i] for i in range(len(ca))}
    db = {tuple(map(int, kb[i])): cb[i] for i in range(len(cb))}
    common = sorted(set(da) & set(db))
    rel = np.array([abs(da[k] - db[k]) / abs(db[k]) for k in common if db[k] != 0])
    only = [k for k in set(da) ^ set(db)]
    worst_only = max((abs(da.get(k, db.get(k))) for k in only), default=0.0)
    W = dev.x.shape[1]
    d_expect = abs(pyffo.pauli_expect(dev, [0] * W) - pyffo.pauli_expect(ref, [0] * W))
    return d_expect, float(rel.max()), float(np.median(rel)), len(only), worst_only


@pytest.mark.parametrize("prec,max_d,max_rel", [(_abi.FP32C, 1e-15, 1e-11), (_abi.FP32, 1e-7, 1e-4)])
@pytest.mark.parametrize("batched", [True, False])
def test_c3_precision_errors(ctx, c2, c2_oracle, prec, max_d, max_rel, batched):
    """FP32 coefficients (plain, or with a stored compensation term) against FP64.
    A key can only be on one side where its |c| is at the threshold's rounding
    distance (mincoeff 1e-5, float spacing 1e-12 there)."""
    gates, x, z, c = c2
    ref = c2_oracle[batched]
    dev = ctx.pauli_propagate(gates, 2, x, z, c, batched=batched, precision=prec, **KI_OPTS)
    d, rmax, rmed, n_only, worst_only = precision_errors(dev, ref)
    print(f"kicked Ising 10x10x20 prec={prec} batched={batched}: |d<O>| {d:.3g}, rel err max {rmax:.3g} median {rmed:.3g}, "
          f"{n_only} keys on one side (largest |c| {worst_only:.3g}) of {len(ref)}", flush=True)
    assert d <= max_d
    assert rmax <= max_rel
    assert worst_only <= KI_OPTS["mincoeff"] * (1 + 1e-5)


@pytest.fixture(scope="module")
def c4():
    inst = wl.batch_kicked_ising(512)
    Wb = max(_abi.words_for_bits(L * L) for (_, _, _, _, L, _) in inst)
    assert Wb == 3 and any(L == 12 for *_, L, _ in inst) and any(L == 11 for *_, L, _ in inst)
    return inst, Wb


@pytest.fixture(scope="module")
def c4_oracle(c4):
    inst, _ = c4

    def one(it):
        g, xi, zi, ci, L, d = it
        W = xi.shape[1]
        r = pyffo.pauli_propagate(g, W, xi, zi, ci, maxdegree=8, mincoeff=1e-4, batched=False)
        return pyffo.pauli_expect(r, [0] * W), len(r)

    with ThreadPoolExecutor(THREADS) as ex:  # the oracle's calls release the GIL
        res = list(ex.map(one, inst))
    return np.array([r[0] for r in res]), np.array([r[1] for r in res])


@pytest.mark.parametrize("layout", ["groups", "windows", "workgroup"])
@pytest.mark.parametrize("prec,tol", [(_abi.FP64, 1e-13), (_abi.FP32C, 1e-12)])
def test_c4_all_512_instances(ctx, c4, c4_oracle, monkeypatch, prec, tol, layout):
    """Every layout of the batch: one term store for all instances in layer-group
    steps or in window steps, and one workgroup per instance."""
    monkeypatch.setenv("FFX_BATCH_MERGED", "0" if layout == "workgroup" else "1")
    if layout == "windows":
        monkeypatch.setenv("FFX_NO_MLAYER", "1")
    inst, Wb = c4

    def pad(a):
        o = np.zeros((a.shape[0], Wb), np.uint64)
        o[:, : a.shape[1]] = a
        return o

    items = [(g, pad(xi), pad(zi), ci, None) for (g, xi, zi, ci, L, d) in inst]
    ex, nt, fb = ctx.pauli_propagate_batch(items, Wb, maxdegree=8, mincoeff=1e-4, batched=False, precision=prec)
    want_e, want_n = c4_oracle
    d = np.abs(ex - want_e)
    bad_n = np.nonzero(nt != want_n)[0]
    print(f"512-instance batch {layout} prec={prec}: max |d<O>| {d.max():.3g}, term-count mismatches {len(bad_n)}, "
          f"fallbacks {fb}", flush=True)
    assert d.max() <= tol, int(np.argmax(d))
    if prec == _abi.FP64:
        assert len(bad_n) == 0, bad_n[:10]
    else:
        # a compensated float can flip a key whose |c| is at the threshold
        assert len(bad_n) <= 2 and all(abs(int(nt[i]) - int(want_n[i])) <= 2 for i in bad_n)


@pytest.mark.parametrize("order", ["bonds", "layers"])
def test_c5_fermi_hubbard_6x6_bit_exact(ctx, order):
    """Both Trotter term orders: bond order (rows of vertical bonds run as disjoint
    layers) and layered order (six disjoint layers per step)."""
    g, (m, c) = wl.fermi_hubbard(6, 6, steps=20, order=order)
    assert len(g) == 6960
    ref = pyffo.majorana_propagate(g, 3, m, c, maxdegree=8, mincoeff=1e-4)
    dev = ctx.majorana_propagate(g, 3, m, c, maxdegree=8, mincoeff=1e-4)
    p = ctx.profile()
    print(f"Fermi-Hubbard 6x6 {order}: {len(dev)} terms, {p['updates']} updates (oracle {ref.updates}), "
          f"{p['run_launches']} launches", flush=True)
    assert_same_terms(dev, ref, exact=True, what=f"Fermi-Hubbard 6x6 {order}")
    assert p["updates"] == ref.updates
    occ = sum(1 << (2 * s) for s in range(36) if s % 2 == 0)
    occw = [occ & ((1 << 64) - 1), occ >> 64]
    assert abs(ctx.expect_last(occw) - pyffo.majorana_expect(ref, occw)) < 1e-15


def test_c5_learned_pruner_kicked_ising_30_layers(ctx):
    """The pruner trained on logs of seven 10x10 circuits of the family prunes the
    30-layer theta in [0.3, 0.7] run every layer; its <Z_55> stays within 1e-3 of
    the oracle's plain FP64 propagation (reference schedule)."""
    L, nq, obs_site = 10, 100, 55
    x, z, c = wl.single_pauli("Z", [obs_site], nq)
    row, col = wl.lattice_geometry(L)
    g5, _ = wl.kicked_ising_family(L)
    rpl = len(g5) // 30
    po = prune_opts(row, col, obs_site, rot_per_layer=rpl)
    circ = [wl.kicked_ising(L, 20 if s % 2 == 0 else 30, theta_lo=0.2 + 0.1 * (s % 2), theta_hi=0.6 + 0.1 * (s % 2),
                            seed=100 + s)[0] for s in range(7)]
    F, T = pruner.pauli_logs(ctx, circ, 2, x, z, c, po, log_every=20, maxdegree=8, mincoeff=1e-5, batched=False)
    ctx.pruner_init(3)
    ctx.pruner_train(F, T, epochs=5, batch=4096, lr=1e-3, seed=1, under_weight=32.0)
    ref = pyffo.pauli_propagate(g5, 2, x, z, c, batched=True, strategy=pyffo.SHARDED, n_threads=THREADS, **KI_OPTS)
    want = pyffo.pauli_expect(ref, [0, 0])
    pl = prune_opts(row, col, obs_site, clifford_tol=1e-12, rot_per_layer=rpl, prune_period=rpl, log10_tau=-6.0)
    _, e = ctx.pauli_propagate_pruned(g5, 2, x, z, c, pl, maxdegree=8, mincoeff=1e-5, batched=False,
                                      precision=_abi.FP32C, download=False, want_expect=True)
    print(f"learned pruner, kicked Ising 10x10x30: |d<Z>| {abs(e - want):.3g} ({len(F)} log rows)", flush=True)
    assert abs(e - want) <= 1e-3


def test_c5_learned_pruner_fermi_hubbard(ctx):
    """The pruner trained on Fermi-Hubbard logs at U = 3 and U = 5 prunes the U = 4
    circuit every Trotter step; <n_0up> stays within 1e-3 of the serial oracle."""
    fh, (fm, fc) = wl.fermi_hubbard(6, 6, steps=20, order="layers")
    frow, fcol = wl.hubbard_geometry(6, 6)
    frpl = len(fh) // 20
    occ = sum(1 << (2 * s) for s in range(36) if s % 2 == 0)
    occw = [occ & ((1 << 64) - 1), occ >> 64]
    fpo = prune_opts(frow, fcol, 0, rot_per_layer=frpl)
    train = [wl.fermi_hubbard(6, 6, u=u, steps=20, order="layers")[0] for u in (3.0, 5.0)]
    F, T = pruner.majorana_logs(ctx, train, 3, fm, fc, fpo,
                                log_every=116, max_rows=1_500_000, maxdegree=8, mincoeff=1e-4, occ=occw)
    ctx.pruner_init(5)
    ctx.pruner_train(F, T, epochs=4, batch=4096, lr=1e-3, seed=2, under_weight=32.0)
    ref = pyffo.majorana_propagate(fh, 3, fm, fc, maxdegree=8, mincoeff=1e-4)
    want = pyffo.majorana_expect(ref, occw)
    pf = prune_opts(frow, fcol, 0, rot_per_layer=frpl, prune_period=frpl, log10_tau=-7.0)
    _, e = ctx.majorana_propagate_pruned(fh, 3, fm, fc, pf, maxdegree=8, mincoeff=1e-4, precision=_abi.FP32C,
                                         download=False, occ=occw)
    print(f"learned pruner, Fermi-Hubbard 6x6: |d<n>| {abs(e - want):.3g} ({len(F)} log rows)", flush=True)
    assert abs(e - want) <= 1e-3
