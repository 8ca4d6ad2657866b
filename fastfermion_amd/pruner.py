"""Learned term pruner: contribution logs -> device training -> pruned propagation.

The pruner has no counterpart in the reference (its truncation is degree, |c|,
top-k and X/Y weight, src/pauli/truncate.h:16-57).  Everything here drives the
device through the C ABI (fastfermion_amd._native.Context):

    rows = pauli_logs(ctx, circuits, ...)        # ffx_pauli_contrib_log per circuit
    ctx.pruner_init(seed); ctx.pruner_train(*rows, under_weight=32)
    ctx.pauli_propagate_pruned(..., prune_opts(..., prune_period=P, log10_tau=tau))

A log row is (64 features, target), target = clamp(log10(|v| + 1e-12), -12, 0)/4
where v is the term's influence on the final expectation; feature 6 is
log10|c|/8, so a row's contribution to <O> has magnitude 10^(8 f6 + 4 t).
"""
from __future__ import annotations

import numpy as np

from ._native import prune_opts  # noqa: F401  (re-exported for callers)


def pauli_logs(ctx, circuits, n_words, x, z, c, po, log_every=20, max_rows=4_000_000, **trunc):
    """Concatenated contribution logs of several Pauli circuits (same observable)."""
    F, T = [], []
    for gates in circuits:
        f, t, _ = ctx.pauli_contrib_log(gates, n_words, x, z, c, po, log_every=log_every, max_rows=max_rows, **trunc)
        F.append(f)
        T.append(t)
    return np.concatenate(F), np.concatenate(T)


def majorana_logs(ctx, circuits, n_words, m, c, po, log_every=20, max_rows=4_000_000, **trunc):
    F, T = [], []
    for gates in circuits:
        f, t, _ = ctx.majorana_contrib_log(gates, n_words, m, c, po, log_every=log_every, max_rows=max_rows,
                                           **trunc)
        F.append(f)
        T.append(t)
    return np.concatenate(F), np.concatenate(T)


def removed_share(feats, targets, pred, frac):
    """Share of the total contribution magnitude sum |c v| carried by the fraction
    `frac` of rows with the lowest score: (by |c| alone, by |c| times the predicted
    influence).  The lower the second number, the better the network ranks terms."""
    logc = feats[:, 6].astype(np.float64) * 8.0
    contrib = 10.0 ** (logc + 4.0 * targets.astype(np.float64))
    total = contrib.sum()
    k = int(frac * len(feats))
    plain = np.argsort(logc, kind="stable")[:k]
    learned = np.argsort(logc + 4.0 * pred.astype(np.float64), kind="stable")[:k]
    return float(contrib[plain].sum() / total), float(contrib[learned].sum() / total)
