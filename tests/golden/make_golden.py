#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference itself.

Runs only where the reference sources exist and ``make -C oracle ref`` has
compiled them into ``oracle/_ref`` (the reference's masterEquation.c,
matgen.c, solv.c, matvec.c, io.c, graphlib.c built by gcc from where they
lie).  The fixtures written here are data only: inputs and the outputs the
reference produced for them.

    python tests/golden/make_golden.py [case ...]

Outputs
  csr_n{n}_{tag}.npz     rowPtr, colInd, val from the reference's gencrs
  cli/{case}/            the reference binary's output files for one CLI run
                         (+ cmd.json with the argv and the kept stdout lines)
"""
from __future__ import annotations

import ctypes as C
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle  # noqa: E402

def rates_hom(n, a, b, g=1.0):
    return np.array([a] + [g] * (n - 1) + [b], np.float64)


def rates_seeded(n, seed, lo=0.1, hi=2.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, n + 1)


def _silence_stdout():
    """gencrs prints one progress line per row; send fd 1 to /dev/null."""
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    os.close(devnull)
    return saved


def _restore_stdout(saved):
    C.CDLL(None).fflush(None)
    os.dup2(saved, 1)
    os.close(saved)


def write_csr_fixtures():
    R = oracle.ref_lib()
    cases = [(3, "a3b7", rates_hom(3, 0.3, 0.7)),
             (4, "seed4", rates_seeded(4, 4)),
             (5, "seed5", rates_seeded(5, 5)),
             (8, "seed8", rates_seeded(8, 8)),
             (10, "cfg1", rates_hom(10, 0.5, 0.5)),
             (10, "seed10", rates_seeded(10, 10))]
    for n, tag, r in cases:
        dim = 1 << n
        nnz = 2 * dim + (1 << (n - 2)) * (n - 1)
        rowPtr = np.zeros(dim + 1, np.uint64)
        colInd = np.zeros(nnz, np.uint64)
        val = np.zeros(nnz, np.float64)
        bit = np.zeros(n, np.int32)
        saved = _silence_stdout()
        try:
            R.gencrs(rowPtr, val, colInd, np.ascontiguousarray(r), bit, dim, n)
        finally:
            _restore_stdout(saved)
        np.savez_compressed(os.path.join(HERE, f"csr_n{n}_{tag}.npz"),
                            n=n, rates=r, rowPtr=rowPtr, colInd=colInd, val=val)
        print(f"csr n={n} {tag}: nnz={int(rowPtr[-1])}")


def _write_vec(path, v):
    with open(path, "w") as f:
        f.write(" ".join("%.17g" % x for x in v))


# Each case: name, n, argv (after the binary), extra files {name: content}.
# Every case keeps the default --save-select: a selection without 'p' leaves
# the reference's `size` uninitialised (src/masterEquation.c:1017,1122), which
# is undefined behaviour, not a fixture.
def cli_cases():
    cases = []
    cases.append(dict(name="n4_ab_f", n=4,
                      argv=["-n", "4", "--alpha", "0.3", "--beta", "0.7",
                            "-t", "0", "-T", "1", "-s", "100", "-f",
                            "-o", "out"], files={}))
    r6 = rates_seeded(6, 6)
    cases.append(dict(name="n6_seed_mf", n=6,
                      argv=["-n", "6", "-i", "rates.txt", "-t", "0", "-T", "2",
                            "-s", "200", "-o", "out"],
                      files={"rates.txt": r6}))
    cases.append(dict(name="n6_seed_f", n=6,
                      argv=["-n", "6", "-i", "rates.txt", "-t", "0", "-T", "2",
                            "-s", "200", "-f", "-o", "out"],
                      files={"rates.txt": r6}))
    # Config-1 rates: the uniform start is exactly stationary -> stops at 100.
    cases.append(dict(name="n10_cfg1_f", n=10,
                      argv=["-n", "10", "-i", "rates.txt", "-t", "0", "-T", "100",
                            "-s", "10000", "-f", "-o", "out"],
                      files={"rates.txt": rates_hom(10, 0.5, 0.5)}))
    # Converging run with inhomogeneous rates (early stop after a transient).
    r8 = rates_seeded(8, 20)
    cases.append(dict(name="n8_seed_conv_f", n=8,
                      argv=["-n", "8", "-i", "rates.txt", "-t", "0", "-T", "200",
                            "-s", "20000", "-f", "-o", "out"],
                      files={"rates.txt": r8}))
    cases.append(dict(name="n8_seed_conv_mf", n=8,
                      argv=["-n", "8", "-i", "rates.txt", "-t", "0", "-T", "200",
                            "-s", "20000", "-o", "out"],
                      files={"rates.txt": r8}))
    # Time-variant rates (sin with amplitude/period/constant files).
    rng = np.random.default_rng(5)
    n = 5
    cases.append(dict(name="n5_tv", n=n,
                      argv=["-n", "5", "-i", "rates.txt", "--time-var",
                            "--amplitude", "amp.txt", "--period", "per.txt",
                            "--constant", "const.txt", "-t", "0", "-T", "3",
                            "-s", "300", "-o", "out"],
                      files={"rates.txt": rates_seeded(5, 55),
                             "amp.txt": rng.uniform(0, 0.5, n + 1),
                             "per.txt": rng.uniform(0.5, 2.0, n + 1),
                             "const.txt": np.ones(n + 1)}))
    # -a snapshots (every 10th step) with the default selection (e, p, c, m):
    # timed expected values, snapshot probabilities, the covariance and
    # multiplicative expected-value directories with per-snapshot PPMs.
    r6a = rates_seeded(6, 66)
    for tag, extra in (("f", ["-f"]), ("mf", [])):
        cases.append(dict(name=f"n6_anim_{tag}", n=6,
                          argv=["-n", "6", "-i", "rates.txt", "-t", "0", "-T", "1",
                                "-s", "100", "-a", "10"] + extra + ["-o", "out"],
                          files={"rates.txt": r6a}))
    rng = np.random.default_rng(57)
    cases.append(dict(name="n5_anim_tv", n=5,
                      argv=["-n", "5", "-i", "rates.txt", "--time-var",
                            "--amplitude", "amp.txt", "--period", "per.txt",
                            "--constant", "const.txt", "-t", "0", "-T", "2",
                            "-s", "200", "-a", "25", "-o", "out"],
                      files={"rates.txt": rates_seeded(5, 58),
                             "amp.txt": rng.uniform(0, 0.5, n + 1),
                             "per.txt": rng.uniform(0.5, 2.0, n + 1),
                             "const.txt": np.ones(n + 1)}))
    return cases


KEEP_PREFIX = ("Generated inital vector", "Real Time elapsed",
               "Simulation time elapsed", "Convergence after", "Confidience",
               "Checksum over result vector", "Number of threads")


def run_cli_case(case):
    d = os.path.join(HERE, "cli", case["name"])
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(d)
    for fname, vec in case["files"].items():
        _write_vec(os.path.join(d, fname), vec)
    p = subprocess.run([oracle.REF_BIN_PATH] + case["argv"], cwd=d,
                       capture_output=True, text=True, stdin=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    lines = []
    for chunk in p.stdout.replace("\r", "\n").split("\n"):
        if chunk.startswith(KEEP_PREFIX) and not chunk.startswith("Real Time"):
            lines.append(chunk)
    # drop per-step progress lines, keep only the final summary ones
    summary = [ln for ln in lines if not ln.startswith("Simulation time elapsed: ")
               or "------" not in ln]
    for junk in ("runtimes.txt",):
        jp = os.path.join(d, junk)
        if os.path.exists(jp):
            os.remove(jp)
    meta = dict(argv=case["argv"], n=case["n"], returncode=p.returncode,
                stdout_lines=summary, stderr=p.stderr.strip().splitlines())
    with open(os.path.join(d, "cmd.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"cli {case['name']}: rc={p.returncode} files={sorted(os.listdir(d))}")


def main():
    if not oracle.have_ref() or not os.path.exists(oracle.REF_BIN_PATH):
        sys.exit("oracle/_ref missing: run `make -C oracle ref` first")
    only = set(sys.argv[1:])        # names of CLI cases to (re)write; none: all
    if not only:
        write_csr_fixtures()
    for case in cli_cases():
        if not only or case["name"] in only:
            run_cli_case(case)


if __name__ == "__main__":
    main()
