"""Spectral bounds of S^-1 H on the device (the GPU edition's two-start
generalized Ritz check, kubo_msd_gpu/src/kubo_msd/metric/gpu_bounds.py:6-74,
with the margins of src/kubo_msd/metric/bounds.py:26-58).

Method: Lanczos for A = S^-1 H, which is self-adjoint in <x, y>_S = x^H S y,
started from the reference's two seeded vectors (NumPy default_rng(811) and
(812) normals), both seeds as the two columns of one complex128 block so every
operator application serves both runs.  Per step one engine action S^-1 H V
(a fused product for S = I, a PCG solve otherwise) and two S inner products:
    u = A v_j - alpha_j v_j - beta_{j-1} v_{j-1},  alpha_j = <v_j, A v_j>_S,
    beta_j = |u|_S,  v_{j+1} = u / beta_j.
The basis is not stored (memory O(N), not O(N x steps)): every `check` steps
the host diagonalises the tridiagonal matrices and estimates the extremal
Ritz residuals |beta_j s_j|; once they are small a second pass replays the
same (deterministic) recurrence accumulating the two extremal Ritz vectors and
measures the true residuals |H x - theta S x| / |x|, which decide convergence
against 1e-10 (|H|_inf + |theta| |S|_inf) as in the GPU edition.  Without
reorthogonalization duplicate Ritz values can appear; extremal ones are not
affected.
"""
import numpy as np
import scipy.linalg as la
import scipy.sparse as sp

from ..errors import NumericalError

SEEDS = (811, 812)


def _row_norm(M):
    return float(np.asarray(abs(sp.csr_matrix(M)).sum(axis=1)).max())


def _start_block(n, seeds):
    return np.stack([np.random.default_rng(s).normal(size=n) for s in seeds], 1).astype(np.complex128)


def _sinner(ops, A, B):
    return ops.inner(A, B).real


class _Lanczos:
    """Stateful Lanczos run on an S-normalised block (one run per column)."""

    def __init__(self, ops, V0, tiny):
        import torch
        self.ops = ops
        self.V = V0
        self.Vprev = torch.zeros_like(V0)
        self.bprev = torch.zeros(V0.shape[1], dtype=torch.float64, device=V0.device)
        self.alphas, self.betas = [], []
        self.tiny = tiny
        self.broken = False

    def advance(self, count, on_vector=None):
        import torch
        for _ in range(count):
            if self.broken:
                return
            if on_vector is not None:
                on_vector(len(self.alphas), self.V)
            U = self.ops.apply_h_action(self.V)
            a = _sinner(self.ops, self.V, U)
            U = U - a.to(U.dtype)[None, :] * self.V - self.bprev.to(U.dtype)[None, :] * self.Vprev
            b = torch.sqrt(torch.clamp(_sinner(self.ops, U, U), min=0.0))
            self.alphas.append(a)
            self.betas.append(b)
            bh = self.ops.ctx.to_host(b, status=True)
            if np.any(bh <= self.tiny):
                # invariant subspace in some column: its Ritz values are exact
                self.broken = True
                return
            self.Vprev, self.V = self.V, U / b.to(U.dtype)[None, :]
            self.bprev = b

    def coefficients(self):
        import torch
        ctx = self.ops.ctx
        return (ctx.to_host(torch.stack(self.alphas), status=True),
                ctx.to_host(torch.stack(self.betas), status=True))


def _ritz(alpha, beta):
    """Extremal Ritz values, their eigenvector components and the Lanczos
    residual estimates |beta_last s_last| for one column."""
    k = len(alpha)
    if k == 1:
        return np.array([alpha[0], alpha[0]]), np.ones((1, 2)), np.array([beta[0], beta[0]])
    theta, S = la.eigh_tridiagonal(alpha, beta[:-1])
    pick = [0, k - 1]
    return theta[pick], S[:, pick], np.abs(beta[-1] * S[-1, pick])


def ritz_extrema(ops, seeds=SEEDS, max_steps=512, check=16, tol=1e-10):
    """Two-start extremal generalized Ritz values of (H, S) with true
    residuals.  Returns dict(extrema (nseed, 2), residuals (nseed, 2),
    steps)."""
    import torch
    n = ops.n
    ns = len(seeds)
    hnorm, snorm = _row_norm(ops.host_H), _row_norm(ops.host_S)
    cap = min(n, max_steps)
    V0 = ops.asarray(_start_block(n, seeds), torch.complex128)
    nrm = torch.sqrt(_sinner(ops, V0, V0))
    if not ops.ctx.status_value(torch.all(nrm > 0)):
        raise NumericalError("empty generalized Ritz subspace")
    V0 = V0 / nrm.to(V0.dtype)[None, :]
    tiny = 1e-14 * max(hnorm, 1e-300)
    run = _Lanczos(ops, V0, tiny)
    while True:
        run.advance(min(check, cap - len(run.alphas)))
        alpha, beta = run.coefficients()
        steps = len(alpha)
        thetas, vecs, est = [], [], []
        for c in range(ns):
            th, S, e = _ritz(alpha[:, c], beta[:, c])
            thetas.append(th)
            vecs.append(S)
            est.append(np.zeros(2) if run.broken else e)
        scale = np.array([[hnorm + abs(t) * snorm for t in th] for th in thetas])
        done = run.broken or steps >= cap
        if not (np.all(np.array(est) <= tol * scale) or done):
            continue
        # second pass: Ritz vectors x = sum_j s_j v_j and their true residuals
        X = torch.zeros((n, 2 * ns), dtype=torch.complex128, device=V0.device)
        coef = np.concatenate(vecs, axis=1)  # (steps, 2*ns)
        coef_t = torch.from_numpy(coef).to(V0.device).to(torch.complex128)

        def accumulate(j, V):
            X.add_(torch.repeat_interleave(V, 2, dim=1) * coef_t[j][None, :])
        _Lanczos(ops, V0, tiny).advance(steps, accumulate)
        theta = torch.from_numpy(np.concatenate(thetas)).to(V0.device)
        R = ops.apply_h_cov(X) - theta.to(X.dtype)[None, :] * ops.S.matmul(X)
        res = ops.ctx.to_host(torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(X, dim=0),
                              status=True)
        res = res.reshape(ns, 2)
        if np.all(np.isfinite(res)) and np.all(res <= tol * scale):
            return dict(extrema=np.array(thetas), residuals=res, steps=steps)
        if done:
            raise NumericalError("generalized Ritz extrema did not converge; enlarge/inspect spectrum and PCG "
                                 "accuracy")


def estimate_spectral_bounds_device(ops, spectral_bounds=None, s_certified=False):
    """The reference's bound report (same keys and margins) from the device
    Ritz extrema: user bounds must enclose the extrema with 10 residuals /
    lambda_min(S); otherwise extrema +- max(1% of the width, that allowance)."""
    r = ritz_extrema(ops)
    lo = float(r["extrema"][:, 0].min())
    hi = float(r["extrema"][:, 1].max())
    residuals = [float(v) for v in r["residuals"].ravel()]
    s_lower = ops.metric.s_lower
    allowance = 10 * max(residuals) / s_lower
    b = _row_norm(ops.host_H) / s_lower
    if spectral_bounds is not None:
        bounds = tuple(float(v) for v in spectral_bounds)
        if bounds[0] >= lo - allowance or bounds[1] <= hi + allowance:
            raise NumericalError("requested spectral bounds do not enclose checked full spectrum with margin")
        certified = bool(s_certified and bounds[0] <= -b and bounds[1] >= b)
        source = "user bounds checked against device two-start generalized Ritz (Lanczos)"
    else:
        if hi - lo <= np.finfo(float).eps * max(abs(lo), abs(hi), 1):
            raise NumericalError("degenerate spectrum needs explicit spectral bounds")
        margin = max(.01 * (hi - lo), allowance)
        bounds = (lo - margin, hi + margin)
        certified = False
        source = "device two-start S-inner-product generalized Ritz (Lanczos) plus margin"
    return {"bounds": bounds, "source": source, "certified": certified, "residuals": residuals,
            "conservative_bounds": (-b, b), "extrema": [lo, hi], "lanczos_steps": r["steps"],
            "seeds": list(SEEDS)}
