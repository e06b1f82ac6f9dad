"""Run steps of one engine configuration and print the time per step.

    python -m tasep_amd.run --n 28 --prec f32 --integ rk4 --steps 5

Rates: alpha, beta ~ U[0.2, 1.0] and bulk ~ LogNormal(0, 0.3) from
--seed (the n=28 precision workload's draw), or --homogeneous A B for
alpha = A, beta = B, bulk 1.  The state starts uniform (2^-n), as the
reference's does.  One warm-up step runs before the timed ones.  Prints one
line per run: n, mode, steps, ms per step (wall time of tasep_run, which
returns after the device has finished) and state-updates/s.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import Engine, OPT_TILED, OPT_XCD_ORDER, OPT_TILE_BLOCKS


def workload_rates(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(0.2, 1.0, 2)
    return np.concatenate([[a], rng.lognormal(0.0, 0.3, n - 1), [b]])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m tasep_amd.run")
    ap.add_argument("--n", type=int, default=28)
    ap.add_argument("--prec", choices=("f32", "f64"), default="f32")
    ap.add_argument("--integ", choices=("rk4", "ref"), default="rk4")
    ap.add_argument("--gen", choices=("mf", "csr"), default="mf")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--h", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=28)
    ap.add_argument("--homogeneous", type=float, nargs=2, metavar=("ALPHA", "BETA"))
    ap.add_argument("--tiled", type=int, default=None, help="TASEP_OPT_TILED")
    ap.add_argument("--xcd-order", type=int, default=None, help="TASEP_OPT_XCD_ORDER")
    ap.add_argument("--tile-blocks", type=int, default=None, help="TASEP_OPT_TILE_BLOCKS")
    args = ap.parse_args(argv)
    n = args.n
    if args.homogeneous:
        a, b = args.homogeneous
        r = np.array([a] + [1.0] * (n - 1) + [b])
    else:
        r = workload_rates(n, args.seed)
    with Engine(n, args.prec, args.integ, args.gen) as e:
        for key, val in ((OPT_TILED, args.tiled), (OPT_XCD_ORDER, args.xcd_order),
                         (OPT_TILE_BLOCKS, args.tile_blocks)):
            if val is not None:
                e.set_option(key, val)
        e.set_rates(r)
        e.set_x0_uniform()
        e.run(0.0, args.h, 1, patience=-1)
        t0 = time.perf_counter()
        info = e.run(0.0, args.h, args.steps, patience=-1, resume=True)
        dt = time.perf_counter() - t0
    ms = dt * 1e3 / max(args.steps, 1)
    print(f"n={n} {args.prec} {args.integ} {args.gen}: {args.steps} steps, {ms:.3f} ms/step, "
          f"{args.steps * (1 << n) / dt:.4e} state-updates/s, mass {info.mass:.12f}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
