"""CPU: the learned-binning oracle pinned against finite differences,
hand-worked features and the bin / decision-rule definitions."""
import math

import numpy as np

from oracle import mlp_oracle as MO


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    P = rng.standard_normal(MO.NPARAM) * 0.3
    X = rng.standard_normal((40, MO.NIN))
    Y = rng.integers(0, MO.NOUT, 40)
    loss, g = MO.loss_and_grad(P, X, Y)
    h = 1e-6
    worst = 0.0
    for p in rng.choice(MO.NPARAM, 120, replace=False):
        e = np.zeros(MO.NPARAM)
        e[p] = h
        fd = (MO.loss_and_grad(P + e, X, Y)[0] - MO.loss_and_grad(P - e, X, Y)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[p]) / max(1e-8, abs(fd) + abs(g[p])))
    assert worst < 1e-5, worst
    # the loss is the mean negative log-likelihood of the labels
    z = MO.forward(P, X)
    ref = np.mean([math.log(np.exp(z[i] - z[i].max()).sum()) + z[i].max() - z[i, Y[i]]
                   for i in range(len(X))])
    assert abs(loss - ref) < 1e-12


def test_features_hand_worked():
    # A (4 x 10): row0 {1,3}, row1 {2}, row2 {}, row3 {0,5,9}
    Ap = [0, 2, 3, 3, 6]
    Ai = [1, 3, 2, 0, 5, 9]
    # output row 0 walks k = 0, 1, 3; output row 1 walks k = 2 (an empty row)
    Lp = [0, 3, 4]
    Li = [0, 1, 3, 2]
    F = MO.row_features(Lp, Li, Ap, Ai, 10)
    l2 = math.log2
    e = 10 * (1 - math.exp(-0.6))  # 6 terms uniform over the span of 10 columns
    # span 10 (columns 0..9), summed single-row spans 3 + 1 + 10 = 14
    want0 = [l2(7), 2, 2, 1, l2(3), l2(11), l2(1 + e), l2(15), l2(7), l2(7), l2(5), 0.75,
             l2(1 + math.sqrt(14)), l2(1 + 14 / 3), 2 - l2(3), l2(1.6)]
    want1 = [0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1.0, 0, 0, 0, 0]
    assert np.allclose(F[0], want0, rtol=0, atol=1e-12)
    assert np.allclose(F[1], want1, rtol=0, atol=1e-12)
    # widths are the rows' column spans: the features do not depend on n
    assert np.array_equal(MO.row_features(Lp, Li, Ap, Ai, 1000), F)


def test_bins_and_decision_rule():
    assert list(MO.bin_of([0, 1, 32, 33, 128, 129, 512, 513, 2048, 2049, 4096, 4097, 10 ** 7])) == \
        [0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5]
    # zero weights: p = softmax(b2) for every row
    P = np.zeros(MO.NPARAM)
    b2 = np.log(np.array([0.5, 0.3, 0.1, 0.05, 0.03, 0.02]))
    P[-6:] = b2
    X = np.zeros((3, MO.NIN))
    assert list(MO.predict_bins(P, X, 0.5)) == [0] * 3
    assert list(MO.predict_bins(P, X, 0.85)) == [2] * 3      # 0.5 + 0.3 + 0.1 = 0.9
    assert list(MO.predict_bins(P, X, 0.96)) == [4] * 3      # ... + 0.05 + 0.03 = 0.98
    assert list(MO.predict_bins(P, X, 1.0)) == [5] * 3
