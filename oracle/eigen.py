"""Independent full-eigenbasis closed forms for the stochastic estimators.
TEST INFRASTRUCTURE ONLY.

Restates the semantics of the reference's benchmarks/metric_eigen_reference.py:39-105
(and the exact solver's pair-analytic forms, src/kubo_msd/solvers/exact.py):
for the generalized problem H C = S C diag(eps) with C^H S C = I and a sample
block B, q = C^H S B.  Then, per sample column b, with Jackson weights
w(E, eps_n) = sum_l G[E, l] T_l((eps_n - c)/a):
  DOS_b(E)            = sum_n w(E, eps_n) |q_nb|^2
  VAC_b(E, t)         = sum_n w(E, eps_n) Re conj(q_nb) e^{+i eps_n t} (V e^{-i eps t} V q)_nb
  MSD commutator_b(E,t)= sum_n w(E, eps_n) |A_nb(t)|^2,
        A(t) = -(V o I(t)) q,  I_nm(t) = t e^{-i(eps_n+eps_m)t/2} sinc((eps_m-eps_n) t / 2)
(all with hbar; V the covariant velocity expressed in the eigenbasis, C^H V C).
These hold for ANY sample block, so every per-sample output is pinned.
"""
import numpy as np
import scipy.linalg as la

HBAR = 0.6582119569509067


def eigenbasis(H, S):
    eps, C = la.eigh(np.asarray(H), np.asarray(S))
    return eps, C


def jackson_weights(eps, E, bounds, M):
    lo, hi = bounds
    c, a = (lo + hi) / 2, (hi - lo) / 2
    m = np.arange(M)
    th = np.pi / (M + 1)
    damping = ((M + 1 - m) * np.cos(m * th) + np.sin(m * th) / np.tan(th)) / (M + 1)
    x = (np.asarray(E) - c) / a
    y = (eps - c) / a
    return ((np.cos(np.arccos(x)[:, None] * m) * damping * np.where(m == 0, 1, 2))
            @ np.cos(m[:, None] * np.arccos(y)[None, :])) / (np.pi * a * np.sqrt(1 - x * x))[:, None]


def time_grid(tmax, dt):
    t = np.arange(int(np.floor(tmax / dt)) + 1) * dt
    if t[-1] < tmax and not np.isclose(t[-1], tmax, rtol=0, atol=8 * np.spacing(max(tmax, dt))):
        t = np.append(t, tmax)
    else:
        t[-1] = tmax
    return t


def observables(H, S, V_axes, B, E, bounds, M, time_fs, integration_dt_fs, stepwise=True):
    """Per-sample exact DOS (nb, ne) and numerators (nb, na, nt, ne)."""
    eps, C = eigenbasis(H, S)
    q = C.conj().T @ (np.asarray(S) @ B)
    w = jackson_weights(eps, E, bounds, M)
    dos = (w @ np.abs(q) ** 2).T
    time_fs = np.asarray(time_fs, float)
    times = np.unique(np.r_[time_grid(time_fs[-1], integration_dt_fs), time_fs])
    sel = np.searchsorted(times, time_fs)
    nb, na, nt, ne = B.shape[1], len(V_axes), len(time_fs), len(E)
    out = {k: np.zeros((nb, na, nt, ne)) for k in
           ("msd_A2/vac_integral", "diffusion_A2_per_fs/vac_integral", "vac_A2_per_fs2/vac_integral")}
    if stepwise:
        out["msd_A2/commutator"] = np.zeros((nb, na, nt, ne))
    for ai, Vcov in enumerate(V_axes):
        V = C.conj().T @ np.asarray(Vcov) @ C
        Vq = V @ q
        vac = np.empty((nb, len(times), ne))
        for k, t in enumerate(times):
            ph = np.exp(-1j * eps * t / HBAR)
            applied = V @ (Vq * ph[:, None])
            terms = (q.conj() * ph.conj()[:, None] * applied).real
            vac[:, k, :] = (w @ terms).T
        Mv = np.zeros((nb, ne))
        Dv = np.zeros((nb, ne))
        for k, t in enumerate(times):
            if k:
                h = t - times[k - 1]
                Mv = Mv + 2 * h * Dv + h * h / 3 * (2 * vac[:, k - 1] + vac[:, k])
                Dv = Dv + h / 2 * (vac[:, k - 1] + vac[:, k])
            hits = np.flatnonzero(sel == k)
            for j in hits:
                out["msd_A2/vac_integral"][:, ai, j] = Mv
                out["diffusion_A2_per_fs/vac_integral"][:, ai, j] = Dv
                out["vac_A2_per_fs2/vac_integral"][:, ai, j] = vac[:, k]
        if stepwise:
            for j, t in enumerate(time_fs):
                ei, ej = eps[:, None], eps[None, :]
                integral = t * np.exp(-.5j * (ei + ej) * t / HBAR) * np.sinc((ej - ei) * t / (2 * np.pi * HBAR))
                amp = -(V * integral) @ q
                out["msd_A2/commutator"][:, ai, j] = (w @ np.abs(amp) ** 2).T
    return dos, out
