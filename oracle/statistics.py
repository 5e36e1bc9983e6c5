"""CPU restatement of the reference's statistics of the stochastic estimator.
Test infrastructure only (the oracle): tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg use it as the checker; the product never imports it.

- block_sums:     solvers/stochastic.py:60-71 (per-sample += into blocks, in
                  sample order)
- msd_gradient:   observables/diffusion.py:4-6
- estimate_ratio: observables/statistics.py:5-23 (paired block bootstrap of
                  sums, numpy.random.default_rng(21491), 256 resamples)
- dos_validity:   solvers/stochastic.py:74-87 (valid / reason masks)
- bootstrap_brute_force: the same estimator by explicit loops over resamples
                  and blocks (an independent statement used to pin the above)

Pinned in tests/test_oracle_statistics.py by the reference's own known
answers (tests/unit/test_statistics.py:4-21: 92/31, zero error for
proportional blocks, no error for one block; tests/unit/test_vac_integrals.py:
4-9: the gradient of 2t^2 is 4t on a nonuniform grid), by exactness of the
second-order gradient on quadratics, and by the brute-force bootstrap.
"""
import numpy as np

LOW_DOS, UNCERTAIN_TRACE, INSUFFICIENT_BLOCKS = 1, 2, 32


def block_sums(per_sample, ids, block_size, nblocks):
    """Sums of per-sample arrays into statistics blocks of block_size
    consecutive sample ids, added one sample at a time in id order."""
    per_sample = np.asarray(per_sample, float)
    out = np.zeros((nblocks,) + per_sample.shape[1:])
    for local, sid in enumerate(ids):
        out[sid // block_size] += per_sample[local]
    return out


def msd_gradient(time_fs, msd, axis=1):
    if len(time_fs) < 2:
        return None
    return .5 * np.gradient(msd, np.asarray(time_fs, float), axis=axis, edge_order=2 if len(time_fs) > 2 else 1)


def resample_weights(nblocks, n_resamples=256, seed=21491):
    """Counts of each block in each resample (rows of n_resamples x nblocks)."""
    rng = np.random.default_rng(seed)
    draws = rng.integers(nblocks, size=(n_resamples, nblocks))
    return np.array([np.bincount(row, minlength=nblocks) for row in draws], float)


def estimate_ratio(A, D, counts, n_resamples=256):
    A = np.asarray(A, float)
    D = np.asarray(D, float)
    counts = np.asarray(counts, int)
    nb = len(counts)
    if A.shape[0] != nb or D.shape[0] != nb or np.any(counts <= 0):
        raise ValueError("inconsistent statistical blocks/counts")
    if D.ndim != 2 or A.shape[-1] != D.shape[-1]:
        if D.shape == (nb, 1):
            D = np.broadcast_to(D, (nb, A.shape[-1]))
        else:
            raise ValueError("statistical denominator must be (blocks,energy)")
    tA, tD = A.sum(0), D.sum(0)
    value = np.full_like(tA, np.nan)
    np.divide(tA, tD, out=value, where=tD != 0)
    if nb < 2:
        return value, None
    w = resample_weights(nb, n_resamples)
    rA = np.tensordot(w, A, axes=(1, 0))
    rD = (w @ D).reshape((n_resamples,) + (1,) * (A.ndim - 2) + (A.shape[-1],))
    ratios = np.full_like(rA, np.nan)
    np.divide(rA, rD, out=ratios, where=rD != 0)
    return value, np.std(ratios, axis=0, ddof=1)


def bootstrap_brute_force(A, D, n_resamples=256):
    """estimate_ratio's standard error by explicit loops: resample r draws
    block b weight[r][b] times; ratio_r = sum of drawn numerators / sum of
    drawn denominators; sample standard deviation over r."""
    A = np.asarray(A, float)
    D = np.asarray(D, float)
    nb = A.shape[0]
    flatA = A.reshape(nb, -1)
    ne = A.shape[-1]
    w = resample_weights(nb, n_resamples)
    out = np.empty(flatA.shape[1])
    for e in range(flatA.shape[1]):
        k = 0 if D.shape[1] == 1 else e % ne
        ratios = []
        for r in range(n_resamples):
            num = sum(w[r, b] * flatA[b, e] for b in range(nb))
            den = sum(w[r, b] * D[b, k] for b in range(nb))
            ratios.append(num / den if den != 0 else np.nan)
        ratios = np.array(ratios)
        mean = sum(ratios) / n_resamples
        out[e] = np.sqrt(sum((x - mean) ** 2 for x in ratios) / (n_resamples - 1))
    return out.reshape(A.shape[1:])


def dos_validity(rho, stderr, threshold, nblocks):
    valid = np.isfinite(rho) & (rho > threshold)
    reason = np.where(valid, 0, LOW_DOS).astype("uint32")
    if stderr is not None:
        uncertain = ~np.isfinite(stderr) | (rho - 1.96 * stderr <= 0)
        valid &= ~uncertain
        reason[uncertain] |= UNCERTAIN_TRACE
    if nblocks < 8:
        reason |= INSUFFICIENT_BLOCKS
    return valid, reason
