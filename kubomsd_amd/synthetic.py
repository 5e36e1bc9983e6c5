"""Seeded synthetic tight-binding models for the KPM transport workloads.

Every generator returns a ``Model`` holding complex128 CSR matrices H and S,
per-axis explicit velocity operators (kind 0, Angstrom/fs) built as
V = (i/hbar) H o D with minimum-image bond displacements D_ij = r_j - r_i, and
the site positions.  For an orthogonal periodic lattice that is exactly the
commutator (i/hbar)[H, X] with X unwrapped across the boundary, the form the
reference requires for periodic cells (a diagonal X from coordinates would
wrap; see the reference's velocity construction, src/kubo_msd/io/model.py:63-72).

Workloads (sizes as used by the benchmarks and tests):
  square_lattice   : 2D square, nearest-neighbour hopping, Anderson disorder
  honeycomb_lattice: 2D honeycomb (two sites per cell), bond length a
  cubic_orbital_model : simple cubic, 4 orbitals/atom, 4x4 blocks to the second
                     shell, nonorthogonal S (BSR-friendly ordering)
"""
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from .units import HBAR_EV_FS

LATTICE_SEED = 20261203


@dataclass
class Model:
    H: sp.csr_matrix
    S: sp.csr_matrix
    velocity: dict
    positions: np.ndarray
    atom_slices: list = field(default_factory=list)
    description: str = ""

    @property
    def n(self):
        return self.H.shape[0]


def _assemble(n, rows, cols, vals):
    m = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    m.sum_duplicates()
    m.sort_indices()
    return m


def _bond_model(n, onsite, rows, cols, hop, disp):
    """H = diag(onsite) + hop on (rows->cols) bonds both ways; V_a = (i/hbar) H o D_a."""
    r = np.concatenate([np.arange(n), rows, cols])
    c = np.concatenate([np.arange(n), cols, rows])
    h = np.concatenate([onsite.astype(np.complex128), hop, np.conj(hop)])
    H = _assemble(n, r, c, h)
    vel = {}
    for a, axis in enumerate("xyz"):
        if disp.shape[1] <= a or not np.any(disp[:, a]):
            continue
        d = np.concatenate([np.zeros(n), disp[:, a], -disp[:, a]])
        vel[axis] = _assemble(n, r, c, 1j * h * d / HBAR_EV_FS)
        vel[axis].eliminate_zeros()
    return H, vel


def square_lattice(L, t=-1.0, W=2.0, a=1.0, periodic=True, seed=LATTICE_SEED + 1):
    """L x L square lattice, hopping t (eV), on-site U[-W/2, W/2] (eV)."""
    rng = np.random.default_rng(seed)
    n = L * L
    onsite = rng.uniform(-W / 2, W / 2, n)
    ix, iy = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    site = ix * L + iy
    rows, cols, disp = [], [], []
    for dx, dy in ((1, 0), (0, 1)):
        jx, jy = ix + dx, iy + dy
        keep = np.ones(n, bool) if periodic else (jx < L) & (jy < L)
        rows.append(site[keep])
        cols.append(((jx % L) * L + (jy % L))[keep])
        disp.append(np.tile([dx * a, dy * a, 0.0], (keep.sum(), 1)))
    rows, cols, disp = np.concatenate(rows), np.concatenate(cols), np.concatenate(disp)
    H, vel = _bond_model(n, onsite, rows, cols, np.full(len(rows), t, np.complex128), disp)
    pos = np.stack([ix * a, iy * a, np.zeros(n)], 1).astype(float)
    return Model(H, sp.identity(n, dtype=np.complex128, format="csr"), vel, pos,
                 description=f"square {L}x{L} t={t} W={W} periodic={periodic}")


def honeycomb_lattice(L1, L2=None, t=-2.7, W=1.0, a=1.42, periodic=True, seed=LATTICE_SEED + 2):
    """Honeycomb L1 x L2 cells (2 sites each), bond length a (Angstrom)."""
    L2 = L1 if L2 is None else L2
    rng = np.random.default_rng(seed)
    ncell = L1 * L2
    n = 2 * ncell
    onsite = rng.uniform(-W / 2, W / 2, n)
    a1 = a * np.array([1.5, np.sqrt(3) / 2, 0.0])
    a2 = a * np.array([1.5, -np.sqrt(3) / 2, 0.0])
    i1, i2 = np.meshgrid(np.arange(L1), np.arange(L2), indexing="ij")
    i1, i2 = i1.ravel(), i2.ravel()
    A = 2 * (i1 * L2 + i2)
    rows, cols, disp = [], [], []
    # B(R + s) sits at R + (a,0,0); A(R) bonds to B(R), B(R - a1), B(R - a2).
    for d1, d2 in ((0, 0), (-1, 0), (0, -1)):
        j1, j2 = i1 + d1, i2 + d2
        keep = np.ones(ncell, bool) if periodic else (j1 >= 0) & (j2 >= 0)
        B = 2 * ((j1 % L1) * L2 + (j2 % L2)) + 1
        delta = d1 * a1 + d2 * a2 + np.array([a, 0.0, 0.0])
        rows.append(A[keep])
        cols.append(B[keep])
        disp.append(np.tile(delta, (keep.sum(), 1)))
    rows, cols, disp = np.concatenate(rows), np.concatenate(cols), np.concatenate(disp)
    H, vel = _bond_model(n, onsite, rows, cols, np.full(len(rows), t, np.complex128), disp)
    R = i1[:, None] * a1 + i2[:, None] * a2
    pos = np.empty((n, 3))
    pos[0::2] = R
    pos[1::2] = R + np.array([a, 0.0, 0.0])
    return Model(H, sp.identity(n, dtype=np.complex128, format="csr"), vel, pos,
                 description=f"honeycomb {L1}x{L2} t={t} W={W} a={a} periodic={periodic}")


def _sym_uniform(rng, count, norb):
    """count symmetric norb x norb matrices with entries U[-1,1]."""
    m = rng.uniform(-1, 1, (count, norb, norb))
    return (m + m.transpose(0, 2, 1)) / 2


def cubic_orbital_model(L, norb=4, a=2.5, jitter=0.05, s_scale=0.08, decay=0.5, onsite_W=6.0,
                        seed=LATTICE_SEED + 3):
    """Simple cubic L^3 atoms, norb orbitals each, blocks to the 2nd neighbour shell.

    H blocks: symmetric U[-1,1] x exp(-(d-a)/decay); on-site diagonal U[-3,3].
    S: identity on-site, off-site s_scale x exp(-(d-a)/decay) x symmetric U[-1,1].
    Orbitals of one atom are contiguous, so H and S are 4x4-block sparse.
    """
    rng = np.random.default_rng(seed)
    nat = L ** 3
    n = nat * norb
    g = np.stack(np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij"), -1).reshape(-1, 3)
    pos = g * a + rng.normal(0, jitter, (nat, 3))
    atom = lambda q: (q[:, 0] % L) * L * L + (q[:, 1] % L) * L + (q[:, 2] % L)
    shells = [v for v in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T
              if 0 < np.abs(v).sum() <= 2]
    half = [v for v in shells if tuple(v) > (0, 0, 0)]  # one of each +/- pair
    hr, hc, hv, sv, dv = [], [], [], [], []
    for v in half:
        i = np.arange(nat)
        j = atom(g + v)
        d = pos[j] + (g + v - g[j]) * a - pos[i]  # minimum-image displacement
        dist = np.linalg.norm(d, axis=1)
        env = np.exp(-(dist - a * np.linalg.norm(v)) / decay)[:, None, None]
        hb = _sym_uniform(rng, nat, norb) * env
        sb = s_scale * _sym_uniform(rng, nat, norb) * env
        hr.append(i)
        hc.append(j)
        hv.append(hb)
        sv.append(sb)
        dv.append(d)
    hr, hc = np.concatenate(hr), np.concatenate(hc)
    hv, sv, dv = np.concatenate(hv), np.concatenate(sv), np.concatenate(dv)
    oo = np.arange(norb)
    def expand(ar, ac, blocks, conj_blocks):
        r = (ar[:, None, None] * norb + oo[None, :, None]).repeat(norb, 2)
        c = (ac[:, None, None] * norb + oo[None, None, :]).repeat(norb, 1)
        rt = (ac[:, None, None] * norb + oo[None, :, None]).repeat(norb, 2)
        ct = (ar[:, None, None] * norb + oo[None, None, :]).repeat(norb, 1)
        return (np.concatenate([r.ravel(), rt.ravel()]), np.concatenate([c.ravel(), ct.ravel()]),
                np.concatenate([blocks.ravel(), conj_blocks.ravel()]))
    onsite = rng.uniform(-onsite_W / 2, onsite_W / 2, n)
    r, c, v = expand(hr, hc, hv, hv.transpose(0, 2, 1))
    H = _assemble(n, np.concatenate([np.arange(n), r]), np.concatenate([np.arange(n), c]),
                  np.concatenate([onsite, v]).astype(np.complex128))
    r, c, v = expand(hr, hc, sv, sv.transpose(0, 2, 1))
    S = _assemble(n, np.concatenate([np.arange(n), r]), np.concatenate([np.arange(n), c]),
                  np.concatenate([np.ones(n), v]).astype(np.complex128))
    vel = {}
    for k, axis in enumerate("xyz"):
        dd = np.broadcast_to(dv[:, k][:, None, None], hv.shape)
        r, c, v = expand(hr, hc, hv * dd, -(hv * dd).transpose(0, 2, 1))
        vel[axis] = _assemble(n, r, c, 1j * v / HBAR_EV_FS)
    slices = [(q * norb, (q + 1) * norb) for q in range(nat)]
    p = np.repeat(pos, norb, axis=0)
    return Model(H, S, vel, p, atom_slices=slices,
                 description=f"simple cubic {L}^3 x {norb} orbitals, 2 shells, nonorthogonal S")


def gershgorin_lower(S):
    """Smallest Gershgorin disc lower bound of a Hermitian S."""
    d = S.diagonal().real
    radii = np.asarray(abs(S).sum(axis=1)).ravel() - np.abs(d)
    return float(np.min(d - radii))


def honeycomb_flake(n_target, W=1.0, t=-2.7, a=1.42, s_scale=0.0, seed=LATTICE_SEED + 5):
    """Open-boundary honeycomb flake with about n_target orbitals; with
    s_scale > 0 a nonorthogonal overlap S = I + s_scale * U[-1,1] on the
    nearest-neighbour bonds (diagonally dominant for s_scale < 1/3)."""
    L = max(2, int(round(np.sqrt(n_target / 2))))
    m = honeycomb_lattice(L, L, t=t, W=W, a=a, periodic=False, seed=seed)
    if s_scale > 0:
        rng = np.random.default_rng(seed + 1)
        off = sp.triu(m.H, k=1).tocoo()
        vals = s_scale * rng.uniform(-1, 1, off.nnz)
        U = sp.coo_matrix((vals, (off.row, off.col)), shape=m.H.shape)
        m.S = (sp.identity(m.n, format="csr") + U + U.T).astype(np.complex128).tocsr()
        m.S.sort_indices()
        m.description += f" nonorthogonal S (bond overlap {s_scale} U[-1,1])"
    return m
