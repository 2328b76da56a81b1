"""Partition quality against a planted/true partition: pair-counting scores as
in include/evaluate.hpp:130-181 (rand index, adjusted rand index, pairwise
precision/recall) and normalised mutual information (arithmetic-mean
normalisation)."""
import numpy as np


def _contingency(truth, pred):
    truth = np.asarray(truth, np.int64)
    pred = np.asarray(pred, np.int64)
    _, t = np.unique(truth, return_inverse=True)
    _, p = np.unique(pred, return_inverse=True)
    table = np.zeros((t.max() + 1, p.max() + 1), np.int64)
    np.add.at(table, (t, p), 1)
    return table


def pairwise_precision_recall(truth, pred):
    """(precision, recall, rand_index, adjusted_rand_index), evaluate.hpp:130-181."""
    c = _contingency(truth, pred).astype(np.float64)
    n = c.sum()
    pairs = n * (n - 1) / 2
    same_both = (c * (c - 1) / 2).sum()
    rows, cols = c.sum(1), c.sum(0)
    same_truth = (rows * (rows - 1) / 2).sum()
    same_pred = (cols * (cols - 1) / 2).sum()
    diff_both = (n * n + (c * c).sum() - (cols * cols).sum() - (rows * rows).sum()) / 2
    ri = (same_both + diff_both) / pairs
    expected = same_truth * same_pred / pairs
    ari = (same_both - expected) / (0.5 * (same_truth + same_pred) - expected)
    precision = same_both / same_pred if same_pred else 1.0
    recall = same_both / same_truth if same_truth else 1.0
    return precision, recall, ri, ari


def nmi(truth, pred):
    """Mutual information / mean of the two entropies (natural log)."""
    c = _contingency(truth, pred).astype(np.float64)
    n = c.sum()
    pij = c / n
    pi, pj = pij.sum(1), pij.sum(0)
    nz = pij > 0
    mi = (pij[nz] * np.log(pij[nz] / np.outer(pi, pj)[nz])).sum()
    hi = -(pi[pi > 0] * np.log(pi[pi > 0])).sum()
    hj = -(pj[pj > 0] * np.log(pj[pj > 0])).sum()
    den = 0.5 * (hi + hj)
    return 1.0 if den == 0 else float(mi / den)
