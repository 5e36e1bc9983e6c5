"""`python -m kubomsd_amd run ...`: the reference's `kubomsd run` flags
(src/kubo_msd/cli.py:10-40) for the stochastic KPM path on the HIP backend,
plus --precision, --device and --device-memory-limit-gb."""
import argparse
import json
from pathlib import Path
import sys

from . import __version__
from .config import RunConfig, resolve_output_grid, resolve_time_grid


def parser():
    p = argparse.ArgumentParser(prog="kubomsd", description="Single-Gamma Kubo MSD transport (MI355X backend); "
                                                            "energies are absolute eV.")
    p.add_argument("--version", action="version", version="kubomsd-amd " + __version__)
    sub = p.add_subparsers(dest="command", required=True)
    run = sub.add_parser("run", help="compute transport from a kubomsd-model-v1 model.json")
    run.add_argument("--input", required=True, type=Path)
    run.add_argument("--out", required=True, type=Path)
    run.add_argument("--solver", choices=["exact", "stochastic-kpm"], default="exact")
    run.add_argument("--backend", choices=["cpu", "hybrid", "gpu", "hip"], default="hip")
    run.add_argument("--kernel", choices=["gaussian", "jackson"], default="gaussian")
    run.add_argument("--stochastic-method", choices=["auto", "stepwise", "spectral"], default="auto")
    run.add_argument("--metric-residual-policy", choices=["fixed", "adaptive"], default="fixed")
    run.add_argument("--metric-check-interval", type=int, default=32)
    run.add_argument("--metric-gap-ratio", type=float, default=.1)
    run.add_argument("--metric-solver", choices=["pcg", "cholesky"], default="pcg")
    run.add_argument("--metric-refine-maxiter", type=int, default=3)
    run.add_argument("--metric-partition-file")
    run.add_argument("--metric-ic-ordering", choices=["rcm", "natural"], default="rcm")
    run.add_argument("--metric-ic-shift-max", type=float, default=1e-2)
    run.add_argument("--metric-gpu-check-batch", type=int, default=1)
    run.add_argument("--cpu-kernel", choices=["cpp", "reference"], default="cpp")
    run.add_argument("--omp-threads", type=int, default=1)
    run.add_argument("--transport-distribution", choices=["serial", "mpi"], default="serial")
    run.add_argument("--emin", type=float)
    run.add_argument("--emax", type=float)
    run.add_argument("--n-energy", type=int, default=201)
    run.add_argument("--energy-grid", type=Path)
    run.add_argument("--time-grid", type=Path)
    run.add_argument("--tmax-fs", type=float)
    run.add_argument("--dt-fs", type=float)
    run.add_argument("--sigma-ev", type=float, default=.05)
    run.add_argument("--n-moments", type=int, default=256)
    run.add_argument("--spectral-min", type=float)
    run.add_argument("--spectral-max", type=float)
    run.add_argument("--axes", default="x,y,z")
    run.add_argument("--axes-vectors", type=Path)
    run.add_argument("--initial-state-selection", choices=["all", "window"], default="all")
    run.add_argument("--buffer-ev", type=float, default=0.)
    for name, default in [("nsamples", 64), ("seed", 20261019), ("sample-batch", 8), ("statistics-block", 8),
                          ("pair-block", 128), ("time-block", 16), ("energy-block", 256), ("metric-maxiter", 1000),
                          ("sampling-max-order", 4096), ("propagation-max-order", 4096), ("device", 0)]:
        run.add_argument("--" + name, type=int, default=default)
    for name, default in [("metric-rtol", 1e-10), ("metric-atol", 0.), ("sampling-tol", 1e-10),
                          ("propagation-tol", 1e-12), ("segment-fs", 1.), ("integration-dt-fs", .01),
                          ("rho-abs", 0.), ("rho-rel", 1e-12), ("memory-limit-gb", 4.),
                          ("device-memory-limit-gb", 0.)]:
        run.add_argument("--" + name, type=float, default=default)
    run.add_argument("--preconditioner", choices=["diagonal", "block-jacobi", "full-cholesky", "cluster-jacobi",
                                                  "ic0"], default="diagonal")
    run.add_argument("--no-vac", action="store_true")
    run.add_argument("--exact-projection-kernel", choices=["sparse", "dense"], default="sparse")
    run.add_argument("--npz", action="store_true", help="accepted for compatibility; results are always NPZ")
    run.add_argument("--efermi-ev", type=float)
    run.add_argument("--efermi-source")
    run.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    run.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    run.add_argument("--block-width-autotune", action="store_true",
                     help="pick the sample batch width from measured step times (sample-batch and its "
                          "doublings up to nsamples that fit the memory budget)")
    run.add_argument("--moment-doubling", action="store_true",
                     help="S = I: moments from the Chebyshev product identities (half the orders; rounding "
                          "differs from the default recursion)")
    sub.add_parser("capabilities", help="report the HIP device and library")
    return p


def config_from_args(args):
    d = vars(args).copy()
    for k in ("command", "input", "out"):
        d.pop(k)
    emin, emax, ne, eg = d.pop("emin"), d.pop("emax"), d.pop("n_energy"), d.pop("energy_grid")
    tg, tm, dt = d.pop("time_grid"), d.pop("tmax_fs"), d.pop("dt_fs")
    lo, hi = d.pop("spectral_min"), d.pop("spectral_max")
    if eg:
        if emin is not None or emax is not None:
            raise ValueError("--energy-grid excludes --emin/--emax")
        E = json.loads(eg.read_text())
    else:
        if emin is None or emax is None:
            raise ValueError("provide --emin and --emax, or --energy-grid")
        E = resolve_output_grid(emin, emax, ne)
    if tg:
        if tm is not None or dt is not None:
            raise ValueError("--time-grid excludes --tmax-fs/--dt-fs")
        T = json.loads(tg.read_text())
    else:
        T = resolve_time_grid(50. if tm is None else tm, .1 if dt is None else dt)
    if (lo is None) != (hi is None):
        raise ValueError("both --spectral-min and --spectral-max required")
    d["spectral_bounds"] = None if lo is None else (lo, hi)
    d["sigma_eV"] = d.pop("sigma_ev")
    d["buffer_eV"] = d.pop("buffer_ev")
    d["efermi_eV"] = d.pop("efermi_ev")
    d["output_npz"] = d.pop("npz")
    d["include_vac"] = not d.pop("no_vac")
    d["autotune"] = not d.pop("no_autotune")
    d["axes"] = tuple(a.strip() for a in d["axes"].split(","))
    if d["axes_vectors"] is not None:
        d["axes_vectors"] = json.loads(d["axes_vectors"].read_text())
    return RunConfig(energy_eV=E, time_fs=T, **d)


def run_input(path, cfg, out):
    from .contracts import RunSummary
    from .errors import KuboMSDError
    from .io.model import load_model
    from .io.results import ResultsWriter
    from .operators.velocity import build_operators
    from .solvers.stochastic import iter_stochastic_blocks
    if cfg.solver != "stochastic-kpm" or cfg.backend != "hip":
        raise KuboMSDError("this backend runs --solver stochastic-kpm on --backend gpu/hip")
    method = "stepwise" if cfg.stochastic_method == "auto" else cfg.stochastic_method
    model = load_model(path)
    ops = build_operators(model, cfg)
    with ResultsWriter(out, cfg) as writer:
        for block in iter_stochastic_blocks(ops, cfg, method):
            writer.append(block)
        summary = RunSummary(str(Path(out).resolve()), "complete", method, "hip",
                             {"h_action_calls": ops.h_calls, "metric": ops.metric.diagnostics()})
        writer.finalize(summary)
    return summary


def main(argv=None):
    p = parser()
    args = p.parse_args(argv)
    if args.command == "capabilities":
        from .backends.hip import build_info
        out = {"version": __version__, "library": build_info()}
        try:
            import torch
            out["device"] = torch.cuda.get_device_name(0) if torch.cuda.is_available() else None
        except Exception as exc:  # report, do not fail
            out["device"] = f"unavailable: {exc}"
        print(json.dumps(out, indent=2))
        return 0
    try:
        cfg = config_from_args(args)
        summary = run_input(args.input, cfg, args.out)
        print(json.dumps({"run_status": summary.run_status, "method": summary.method, "backend": summary.backend,
                          "output": summary.output_path}, indent=2))
        return 0
    except (ValueError, OSError, KeyError, TypeError, MemoryError, RuntimeError) as exc:
        print(f"kubomsd: {type(exc).__name__}: {exc}", file=sys.stderr)
        return 2


if __name__ == "__main__":
    raise SystemExit(main())
