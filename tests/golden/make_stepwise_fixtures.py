"""Writes tests/golden/stepwise_fixtures.npz: per-sample stepwise DOS and MSD
numerators of the FP64 oracle (oracle/kpm.py on oracle/native.py, which is
bit-identical to the reference's compiled CPU kernel, tests/test_oracle.py)
for the first samples of two workloads that the GPU tests run in full:

  c1: square 100x100 periodic (N = 10^4), t = -1 eV, W = 2 eV, bounds
      (-5.1, 5.1), M = 512, E = linspace(-4, 4, 161), t = 0 .. 10 fs every
      0.5 fs, MSD along x; samples 0..3 of seed 20261019.
  hc: honeycomb 100x100 cells periodic (N = 2*10^4), t = -2.7 eV, W = 1 eV,
      bounds (-8.8, 8.8), M = 2048, E = linspace(-8, 8, 161), t = 0, 1, 2,
      3 fs, MSD along x and y; samples 0..1 of seed 20261019.

Samples are the reference's unit-modulus phases exp(2 pi i u), u from
PCG64(SeedSequence([seed, id])) (src/kubo_msd/metric/sampling.py:8-9).  The
model arrays are hashed into the file so a changed generator is detected.
Run from the repository root: python tests/golden/make_stepwise_fixtures.py
(about a minute on 8 cores).
"""
import hashlib
import os
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from kubomsd_amd.synthetic import honeycomb_lattice, square_lattice  # noqa: E402
from oracle import kpm  # noqa: E402

SEED = 20261019
OUT = Path(__file__).resolve().parent / "stepwise_fixtures.npz"


def phases(n, ids):
    return np.column_stack([np.exp(2j * np.pi * np.random.Generator(np.random.PCG64(np.random.SeedSequence(
        [SEED, int(i)]))).random(n)) for i in ids])


def model_hash(m, axes):
    h = hashlib.sha256()
    for M in [m.H] + [m.velocity[a] for a in axes]:
        for arr in (M.indptr, M.indices, M.data):
            h.update(np.ascontiguousarray(arr).tobytes())
    return h.hexdigest()


def run(m, axes, bounds, M, E, T, ids, threads):
    vel = {a: m.velocity[a] for a in axes}
    cfg = kpm.config(energy_eV=E, time_fs=T, spectral_bounds=bounds, n_moments=M, segment_fs=1.0,
                     axes_vectors=np.eye(3)[["xyz".index(a) for a in axes]])
    ops = kpm.Ops(m.H, m.S, vel, rtol=1e-10, threads=threads)
    return kpm.stepwise_samples(ops, cfg, phases(m.n, ids))


def main():
    threads = len(os.sched_getaffinity(0))
    out = {}
    c1 = square_lattice(100, t=-1.0, W=2.0)
    dos, nums = run(c1, "x", (-5.1, 5.1), 512, np.linspace(-4, 4, 161), np.arange(21) * 0.5, range(4), threads)
    out.update(c1_dos=dos, c1_msd=nums["msd_A2/commutator"], c1_ids=np.arange(4), c1_hash=model_hash(c1, "x"))
    hc = honeycomb_lattice(100, 100, t=-2.7, W=1.0)
    dos, nums = run(hc, "xy", (-8.8, 8.8), 2048, np.linspace(-8, 8, 161), np.arange(4.0), range(2), threads)
    out.update(hc_dos=dos, hc_msd=nums["msd_A2/commutator"], hc_ids=np.arange(2), hc_hash=model_hash(hc, "xy"))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
