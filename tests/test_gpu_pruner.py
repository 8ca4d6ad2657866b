"""Pruned mode, contribution logs and the learned pruner on the device.

* MLP forward and loss gradient (MFMA kernels) against the numpy restatement
  with the same bf16 rounding points (oracle/pyprune.py);
* training (device Adam) lowers the loss;
* contribution logs: sum_q c_q(e) v_q(e) equals the final expectation at every
  logged step (an exact identity of the linear map the truncated propagation
  applies), for Pauli and Majorana circuits, batched and per-gate schedules;
  without truncation the influence of the observable equals the dense
  statevector expectation; the logged features of the first step equal the
  numpy features of the observable's terms;
* Clifford rotations (|cos theta| <= tol run as fused key remaps) reproduce the
  plain FP64 result up to the dropped cos branches (6e-17 relative);
* learned pruning with a threshold nothing reaches leaves the result
  bit-identical to the same run without it.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import pyprune
from _dense import basis_vector, pauli_matrix, rot_unitary
from _helpers import assert_close_terms, assert_same_terms
from fastfermion_amd import _abi, workloads as wl
from fastfermion_amd._native import Context, prune_opts

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(pyprune.NPARAMS, np.float32)
    p[pyprune.OFF_W1:pyprune.OFF_B1] = rng.normal(0, math.sqrt(2 / 64), 64 * 128)
    p[pyprune.OFF_B1:pyprune.OFF_W2] = rng.normal(0, 0.1, 128)
    p[pyprune.OFF_W2:pyprune.OFF_B2] = rng.normal(0, math.sqrt(2 / 128), 128 * 128)
    p[pyprune.OFF_B2:pyprune.OFF_W3] = rng.normal(0, 0.1, 128)
    p[pyprune.OFF_W3:pyprune.OFF_B3] = rng.normal(0, math.sqrt(1 / 128), 128)
    p[pyprune.OFF_B3] = -1.0
    return p


def rand_feats(rng, n):
    X = rng.uniform(0, 1, (n, 64)).astype(np.float32)
    X[rng.uniform(0, 1, X.shape) < 0.5] = 0.0
    return pyprune.bf16(X)


@pytest.mark.parametrize("n", [16, 300, 4099])
def test_forward_matches_oracle(ctx, n):
    rng = np.random.default_rng(n)
    p = rand_params(3)
    X = rand_feats(rng, n)
    ctx.pruner_set_params(p)
    y = ctx.pruner_forward(X)
    yo = pyprune.mlp_forward(p, X)
    err = np.abs(y - yo)
    assert err.max() <= 2e-3 * (1.0 + np.abs(yo).max()), err.max()
    assert np.linalg.norm(y - yo) <= 1e-4 * np.linalg.norm(yo) + 1e-3


@pytest.mark.parametrize("n,under", [(64, 1.0), (1000, 1.0), (1000, 4.0)])
def test_loss_grad_matches_oracle(ctx, n, under):
    rng = np.random.default_rng(100 + n)
    p = rand_params(4)
    X = rand_feats(rng, n)
    T = rng.uniform(-3, 0, n).astype(np.float32)
    ctx.pruner_set_params(p)
    loss, g = ctx.pruner_loss_grad(X, T, under_weight=under)
    lo, go = pyprune.mlp_loss_grad(p, X, T, under_weight=under)
    assert abs(loss - lo) <= 1e-4 * lo, (loss, lo)
    blocks = {"W1": (pyprune.OFF_W1, pyprune.OFF_B1), "b1": (pyprune.OFF_B1, pyprune.OFF_W2),
              "W2": (pyprune.OFF_W2, pyprune.OFF_B2), "b2": (pyprune.OFF_B2, pyprune.OFF_W3),
              "w3": (pyprune.OFF_W3, pyprune.OFF_B3), "b3": (pyprune.OFF_B3, pyprune.NPARAMS)}
    for name, (a, b) in blocks.items():
        d = np.linalg.norm(g[a:b] - go[a:b])
        assert d <= 5e-3 * np.linalg.norm(go[a:b]) + 1e-7, (name, d, np.linalg.norm(go[a:b]))


def test_training_lowers_loss(ctx):
    rng = np.random.default_rng(9)
    n = 20000
    X = rand_feats(rng, n)
    w = rng.normal(0, 1, 64)
    T = (np.tanh(X @ w / 8.0) - 1.0).astype(np.float32)
    ctx.pruner_init(11)
    l0, _ = ctx.pruner_loss_grad(X, T)
    hist = ctx.pruner_train(X, T, epochs=6, batch=1024, lr=2e-3, seed=5)
    l1, _ = ctx.pruner_loss_grad(X, T)
    assert np.all(np.isfinite(hist))
    assert l1 < 0.1 * l0, (l0, hist, l1)


def ising_case(L, layers):
    gates, thetas = wl.kicked_ising(L, layers)
    nq = L * L
    centre = (L // 2) * L + L // 2
    x, z, c = wl.single_pauli("Z", [centre], nq)
    row, col = wl.lattice_geometry(L)
    po = prune_opts(row, col, centre, rot_per_layer=len(gates) // layers)
    return gates, _abi.words_for_bits(nq), x, z, c, po


@pytest.mark.parametrize("batched", [False, True])
def test_log_invariant_pauli(ctx, batched):
    gates, W, x, z, c, po = ising_case(4, 5)
    kw = dict(maxdegree=6, mincoeff=1e-4, batched=batched)
    _, ex = ctx.pauli_propagate(gates, W, x, z, c, download=False, want_expect=True, **kw)
    feats, targ, sums = ctx.pauli_contrib_log(gates, W, x, z, c, po, log_every=17, **kw)
    assert len(sums) == len(gates) // 17 + 1
    assert np.all(np.abs(sums - ex) <= 1e-12 + 1e-10 * abs(ex)), (ex, sums[:5])
    assert len(feats) > len(sums) and feats.shape[1] == 64
    assert np.all((targ <= 0) & (targ >= -3))


def test_log_invariant_majorana(ctx):
    gates, (m, c) = wl.fermi_hubbard(2, 2, steps=3)
    W = 1
    row, col = wl.hubbard_geometry(2, 2)
    po = prune_opts(row, col, 0, rot_per_layer=len(gates) // 3)
    occ = [0b0101]
    _, ex = ctx.majorana_propagate(gates, W, m, c, maxdegree=6, mincoeff=1e-5, download=False, occ=occ)
    feats, targ, sums = ctx.majorana_contrib_log(gates, W, m, c, po, log_every=7, maxdegree=6, mincoeff=1e-5,
                                                 occ=occ)
    assert len(sums) == len(gates) // 7 + 1
    assert np.all(np.abs(sums - ex) <= 1e-12 + 1e-10 * abs(ex)), (ex, sums[:5])
    assert len(feats) > 0


def test_log_influence_exact_without_truncation(ctx):
    rng = np.random.default_rng(21)
    n = 5
    spec = []
    for _ in range(40):
        w = int(rng.integers(1, 4))
        qs = sorted(rng.choice(n, w, replace=False).tolist())
        spec.append(("".join("XYZ"[int(rng.integers(3))] for _ in qs), tuple(qs), float(rng.uniform(-1.5, 1.5))))
    gates = wl._pack(spec)
    x, z, c = wl.single_pauli("ZX", [1, 3], n)
    row, col = [0] * n, list(range(n))
    po = prune_opts(row, col, 1, rot_per_layer=8)
    feats, targ, sums = ctx.pauli_contrib_log(gates, 1, x, z, c, po, log_every=5, maxdegree=None, mincoeff=0.0,
                                              batched=False)
    # dense: <0|U^dag O U|0> with U the circuit (first gate applied first)
    U = np.eye(2 ** n, dtype=complex)
    for ops, qs, th in spec:
        letters = ["I"] * n
        for o, q in zip(ops, qs):
            letters[q] = o
        U = rot_unitary("".join(letters), th) @ U
    O = pauli_matrix("IZIXI")
    psi = U @ basis_vector(n, [])
    exact = np.vdot(psi, O @ psi).real
    assert np.all(np.abs(sums - exact) <= 1e-12), (exact, sums)
    # the last rows are the observable's own (logged at step 0): target = log10|v|/4, v = exact / c
    want = max(min(math.log10(abs(exact) + 1e-12), 0.0), -12.0) / 4
    assert abs(targ[-1] - want) <= 1e-6


def test_log_features_of_first_step(ctx):
    L = 4
    gates, thetas = wl.kicked_ising(L, 3)
    W = 1
    spec = [("Z", (5,), 0.7), ("XY", (1, 2), -0.2), ("ZZZ", (0, 9, 15), 0.05)]
    xs, zs, cs = [], [], []
    for ops, qs, cf in spec:
        xx, zz = _abi.pauli_word(ops, qs, W)
        xs.append(xx)
        zs.append(zz)
        cs.append(cf)
    x, z, c = np.array(xs).reshape(-1, W), np.array(zs).reshape(-1, W), np.array(cs, np.complex128)
    row, col = wl.lattice_geometry(L)
    rpl = len(gates) // 3
    po = prune_opts(row, col, 6, rot_per_layer=rpl)
    feats, targ, sums = ctx.pauli_contrib_log(gates, W, x, z, c, po, log_every=10 ** 6, maxdegree=8,
                                              mincoeff=1e-6)
    assert len(sums) == 1 and len(feats) == 3
    # mean |theta| of the branching rotations, summed in circuit order as the host does
    acc, cnt = 0.0, 0
    for t in gates["theta"]:
        if abs(math.cos(t)) > 1e-6:
            acc += abs(math.atan2(math.sin(t), math.cos(t)))
            cnt += 1
    mean_theta = float(np.float32(acc / cnt))
    want = np.array([pyprune.term_features((x[i], z[i]), abs(cs[i]), "pauli", row, col, 6, len(gates) / rpl,
                                           mean_theta, 0.0, W) for i in range(3)])
    order_g = np.lexsort(feats.T[::-1])
    order_w = np.lexsort(want.T[::-1])
    assert np.array_equal(feats[order_g], want[order_w]), np.abs(feats[order_g] - want[order_w]).max()


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("prec", [_abi.FP64, _abi.FP32C])
def test_clifford_rotations_match_plain(ctx, batched, prec):
    gates, W, x, z, c, po = ising_case(5, 6)
    po.clifford_tol = 1e-12
    kw = dict(maxdegree=8, mincoeff=1e-5, batched=batched, precision=prec)
    plain, ex0 = ctx.pauli_propagate(gates, W, x, z, c, want_expect=True, **kw)
    upd0 = ctx.profile()["updates"]
    fast, ex1 = ctx.pauli_propagate_pruned(gates, W, x, z, c, po, want_expect=True, **kw)
    upd1 = ctx.profile()["updates"]
    tol = 1e-13 if prec == _abi.FP64 else 1e-7
    assert_close_terms(fast, plain, tol, mincoeff=1e-5, what="clifford rotations")
    assert abs(ex1 - ex0) <= tol
    if batched:
        assert upd1 < upd0  # the tiny cos copies no longer branch inside the commuting batches
    elif prec == _abi.FP64:
        # per-gate truncation erases the tiny cos copies at once: every rotation sees the same live terms
        assert upd1 == upd0


def xx_clifford_case(rng, nq, layers):
    """Runs of X-type Clifford rotations (XX(+-pi/2), X(pi/2)) between layers of
    generic single-qubit Z/Y rotations: every run's axes commute pairwise."""
    spec = []
    for _ in range(layers):
        for _ in range(int(rng.integers(6, 14))):
            a, b = sorted(rng.choice(nq, 2, replace=False).tolist())
            ax = "XX" if rng.uniform() < 0.8 else "X"
            qs = (a, b) if ax == "XX" else (a,)
            spec.append((ax, qs, float(rng.choice([math.pi / 2, -math.pi / 2]))))
        for q in rng.choice(nq, int(rng.integers(3, 8)), replace=False).tolist():
            spec.append((str(rng.choice(["Z", "Y"])), (q,), float(rng.uniform(-1.0, 1.0))))
    return wl._pack(spec)


@pytest.mark.parametrize("case", ["ising_pergate", "ising_batched", "ising_xwcap", "xx_runs"])
def test_commuting_runs_match_serial_loop(ctx, monkeypatch, case):
    """A run of pairwise-commuting Clifford rotations is applied from the set of
    rotations acting on each term (fixed by its key at the start of the run);
    keys, coefficients and work counters must equal the rotation-by-rotation loop."""
    kw = dict(maxdegree=7, mincoeff=1e-5, batched=case == "ising_batched")
    if case.startswith("ising"):
        gates, W, x, z, c, po = ising_case(5, 6)
        if case == "ising_xwcap":
            kw.update(max_xweight=3, xtrunc_period=4)
    else:
        rng = np.random.default_rng(21)
        nq = 40
        gates = xx_clifford_case(rng, nq, 8)
        W = 1
        x, z, c = wl.single_pauli("Z", [7], nq)
        x2, z2, c2 = wl.single_pauli("Y", [11], nq)
        x, z, c = np.concatenate([x, x2]), np.concatenate([z, z2]), np.concatenate([c, 0.5 * c2])
        row, col = [0] * nq, list(range(nq))
        po = prune_opts(row, col, 7, rot_per_layer=16)
    po.clifford_tol = 1e-12
    res = {}
    for serial in ("1", "0"):
        monkeypatch.setenv("FFX_CROT_SERIAL", serial)
        r = ctx.pauli_propagate_pruned(gates, W, x, z, c, po, **kw)
        res[serial] = (r, ctx.profile())
    assert_same_terms(res["0"][0], res["1"][0], exact=True, what=f"commuting runs {case}")
    for k in ("updates", "outs", "peak_terms"):
        assert res["0"][1][k] == res["1"][1][k], (k, res["0"][1][k], res["1"][1][k])


def test_clifford_rotations_majorana(ctx):
    rng = np.random.default_rng(8)
    nm = 12
    gl = []
    for i in range(60):
        d = 2 if rng.uniform() < 0.6 else 4
        idx = sorted(rng.choice(nm, d, replace=False).tolist())
        th = float(rng.choice([math.pi / 2, -math.pi / 2])) if i % 3 == 0 else float(rng.uniform(-1, 1))
        gl.append((idx, th, (d >> 1) & 1))
    gates = wl.pack_majorana(gl)
    m = np.zeros((1, 1), np.uint64)
    m[0, 0] = np.uint64(0b11)
    c = np.array([1j])
    row, col = [0] * nm, list(range(nm))
    po = prune_opts(row, col, 0, clifford_tol=1e-12)
    plain = ctx.majorana_propagate(gates, 1, m, c, maxdegree=6, mincoeff=1e-6)
    fast = ctx.majorana_propagate_pruned(gates, 1, m, c, po, maxdegree=6, mincoeff=1e-6)
    assert_close_terms(fast, plain, 1e-13, mincoeff=1e-6, what="majorana clifford rotations")


def test_pruning_below_every_score_is_a_noop(ctx):
    gates, W, x, z, c, po = ising_case(5, 4)
    po.clifford_tol = 1e-12
    kw = dict(maxdegree=8, mincoeff=1e-5, batched=False)
    ref = ctx.pauli_propagate_pruned(gates, W, x, z, c, po, **kw)
    ctx.pruner_set_params(rand_params(2))
    po.prune_period = 10
    po.log10_tau = -1.0e30
    got = ctx.pauli_propagate_pruned(gates, W, x, z, c, po, **kw)
    assert_same_terms(got, ref, exact=True, what="pruning with an unreachable threshold")


def test_pruning_erases_terms(ctx):
    gates, W, x, z, c, po = ising_case(5, 4)
    kw = dict(maxdegree=8, mincoeff=1e-5, batched=False)
    ref = ctx.pauli_propagate(gates, W, x, z, c, **kw)
    p = np.zeros(pyprune.NPARAMS, np.float32)
    p[pyprune.OFF_B3] = -0.25  # constant prediction y = -0.25: score log10|c| - 1
    ctx.pruner_set_params(p)
    po.prune_period = 7
    po.log10_tau = -3.0  # erases |c| < 1e-2 every 7th rotation
    got = ctx.pauli_propagate_pruned(gates, W, x, z, c, po, **kw)
    assert 0 < len(got) < len(ref)
    # every survivor of the last pruning point had |c| >= 1e-2 there; the later rotations only
    # scale or branch them, so no survivor is far below the threshold
    assert np.abs(got.c).max() > 1e-2


def test_trained_network_ranks_terms_better_than_coefficients(ctx):
    """What the trained network gains over truncation by |c| alone: on the logs of a
    held-out circuit (seed not in the training set), erasing the half / four fifths
    of the terms with the lowest predicted contribution removes far less of the
    total contribution than erasing those with the lowest |c|."""
    from fastfermion_amd import pruner

    L = 6
    nq, centre = L * L, (L // 2) * L + L // 2
    x, z, c = wl.single_pauli("Z", [centre], nq)
    row, col = wl.lattice_geometry(L)
    po = prune_opts(row, col, centre, rot_per_layer=2 * L * (L - 1) + nq)
    trunc = dict(maxdegree=8, mincoeff=1e-5, batched=False)
    train = [wl.kicked_ising(L, 12, theta_lo=0.3, theta_hi=0.7, seed=s)[0] for s in range(4)]
    F, T = pruner.pauli_logs(ctx, train, 1, x, z, c, po, log_every=10, **trunc)
    Fh, Th = pruner.pauli_logs(ctx, [wl.kicked_ising(L, 12, theta_lo=0.3, theta_hi=0.7, seed=99)[0]], 1, x, z, c, po,
                               log_every=10, **trunc)
    assert len(F) > 20000 and len(Fh) > 2000
    ctx.pruner_init(3)
    hist = ctx.pruner_train(F, T, epochs=5, batch=2048, lr=1e-3, seed=1, under_weight=16.0)
    assert hist[-1] < hist[0]
    y = ctx.pruner_forward(Fh)
    for frac in (0.5, 0.8):
        plain, learned = pruner.removed_share(Fh, Th, y, frac)
        assert learned < 0.2 * plain, (frac, plain, learned)
