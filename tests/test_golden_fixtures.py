"""The committed stepwise fixtures (tests/golden/stepwise_fixtures.npz) are
the oracle's output for today's generators: the model hashes match and the
DOS of the stored samples is recomputed by the oracle chain (bit for bit, on
the same column block; the DOS is one moment chain, cheap on the CPU).  The GPU
accuracy tests compare the device estimator against these fixtures."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_stepwise_fixtures as mk  # noqa: E402
from kubomsd_amd.synthetic import honeycomb_lattice, square_lattice  # noqa: E402
from oracle import kpm  # noqa: E402


def dos(m, bounds, M, E, ids):
    cfg = kpm.config(energy_eV=E, time_fs=[0.], spectral_bounds=bounds, n_moments=M)
    ops = kpm.Ops(m.H, m.S, {}, rtol=1e-10, threads=4)
    return kpm.projected_quadratic(mk.phases(m.n, ids), ops, cfg)


def test_fixture_matches_generators_and_oracle():
    fx = np.load(mk.OUT)
    c1 = square_lattice(100, t=-1.0, W=2.0)
    assert mk.model_hash(c1, "x") == str(fx["c1_hash"])
    np.testing.assert_array_equal(dos(c1, (-5.1, 5.1), 512, np.linspace(-4, 4, 161), fx["c1_ids"]), fx["c1_dos"])
    hc = honeycomb_lattice(100, 100, t=-2.7, W=1.0)
    assert mk.model_hash(hc, "xy") == str(fx["hc_hash"])
    np.testing.assert_array_equal(dos(hc, (-8.8, 8.8), 2048, np.linspace(-8, 8, 161), fx["hc_ids"]), fx["hc_dos"])
    assert fx["c1_msd"].shape == (4, 1, 21, 161) and fx["hc_msd"].shape == (2, 2, 4, 161)
    assert np.all(fx["c1_msd"][:, :, 0] == 0) and np.all(np.abs(fx["c1_msd"][:, :, 1:]).max() > 0)
