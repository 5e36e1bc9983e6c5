"""Batched integration nodes of the spectral estimator (host torch, no
device): _node_vac against the per-node contraction of the reference
(src/kubo_msd/solvers/spectral.py:38-46: q = table beta, mu_l = Re sum_p
1/2 (q_{l+p} + q_{|l-p|}) conj(beta_p), vac = mu G^T), with node orders below
the table's, several node chunks, and _segments against the reference's
anchor rule on an irregular node grid."""
import numpy as np
import torch

from kubomsd_amd.solvers.spectral import _node_vac, _segments


def per_node(table, b, G, L):
    Kt = len(b) - 1
    q = table[:, :, :Kt + 1] @ b
    l = np.arange(L)[:, None]
    p = np.arange(Kt + 1)[None, :]
    mu = np.sum(.5 * (q[:, l + p] + q[:, np.abs(l - p)]) * b.conj()[None, None, :], axis=2).real
    return mu @ G.T


def test_node_batch_matches_per_node_contraction():
    rng = np.random.default_rng(3)
    nb, L, K, ne = 3, 40, 9, 7
    table = rng.normal(size=(nb, L + K, K + 1)) + 1j * rng.normal(size=(nb, L + K, K + 1))
    G = rng.normal(size=(ne, L))
    orders = [0, 1, 4, 9, 9, 6, 2]
    beta = np.zeros((K + 1, len(orders)), complex)
    want = []
    for j, k in enumerate(orders):
        b = rng.normal(size=k + 1) + 1j * rng.normal(size=k + 1)
        beta[:k + 1, j] = b
        want.append(per_node(table, b, G, L))
    for chunk in (1, 16 * nb * L * (K + 1) * 3 * 2, 1 << 30):  # one node, two nodes, all nodes per chunk
        got = _node_vac(torch.from_numpy(table), torch.from_numpy(beta), torch.from_numpy(G), L, chunk).numpy()
        np.testing.assert_allclose(got, np.array(want), rtol=1e-12, atol=1e-12 * np.abs(want).max())


def reference_anchors(times, segment_fs):
    """The reference's loop (spectral.py:24-36): the anchor each node's tables
    were built at."""
    out, tables, anchor, next_anchor, prev_t = [], None, 0., -1., 0.
    for t in times:
        if tables is None or t > next_anchor + 1e-14:
            anchor = 0. if tables is None else float(prev_t)
            next_anchor = min(max(anchor + segment_fs, float(t)), float(times[-1]))
            tables = True
        out.append(anchor)
        prev_t = t
    return out


def test_segments_follow_reference_anchor_rule():
    times = np.unique(np.r_[np.arange(0, 2.05, 0.1), [0.33, 1.01, 1.999]])
    for seg in (0.25, 0.5, 1.0, 5.0):
        anchors = [None] * len(times)
        for anchor, end, nodes in _segments(times, seg):
            for i in nodes:
                anchors[i] = anchor
                assert times[i] <= end + 1e-14
        assert anchors == reference_anchors(times, seg)
