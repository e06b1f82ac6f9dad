"""Readers for the committed golden fixtures (tests/golden/)."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def csr_fixtures():
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, "csr_n*.npz"))):
        z = np.load(p)  # allow_pickle=False by default
        out.append(dict(name=os.path.basename(p), n=int(z["n"]), rates=z["rates"],
                        rowPtr=z["rowPtr"], colInd=z["colInd"], val=z["val"]))
    return out


def _read_vec(path):
    with open(path) as f:
        return np.array([float(v) for v in f.read().replace(",", " ").split()],
                        np.float64)


def cli_case_names():
    return sorted(os.path.basename(d) for d in glob.glob(os.path.join(GOLDEN, "cli", "*"))
                  if os.path.isdir(d))


def cli_case(name):
    """Parsed inputs + expected outputs of one reference CLI run."""
    d = os.path.join(GOLDEN, "cli", name)
    with open(os.path.join(d, "cmd.json")) as f:
        meta = json.load(f)
    argv = meta["argv"]
    opts, flags = {}, set()
    i = 0
    while i < len(argv):
        a = argv[i]
        if a in ("-f", "--time-var"):
            flags.add(a)
            i += 1
        else:
            opts[a] = argv[i + 1]
            i += 2
    n = int(opts["-n"])
    if "-i" in opts:
        rates = _read_vec(os.path.join(d, opts["-i"]))[: n + 1]
    else:
        a, b = float(opts["--alpha"]), float(opts["--beta"])
        g = float(opts.get("--gamma", 1.0))
        rates = np.array([a] + [g] * (n - 1) + [b])
    case = dict(name=name, dir=d, n=n, rates=rates, t0=float(opts["-t"]),
                T=float(opts["-T"]), steps=int(opts["-s"]), csr="-f" in flags,
                tv=None, meta=meta, argv=argv)
    if "--time-var" in flags:
        def opt_vec(key, default):
            if key in opts:
                return _read_vec(os.path.join(d, opts[key]))[: n + 1]
            return np.full(n + 1, default, np.float64)
        case["tv"] = dict(type=np.zeros(n + 1, np.int32),
                          amp=opt_vec("--amplitude", 1.0),
                          period=opt_vec("--period", 1.0),
                          offset=opt_vec("--offset", 0.0),
                          c=opt_vec("--constant", 1.0))
    # expected outputs
    probs = os.path.join(d, "out")
    if os.path.exists(probs):
        xs, bits = [], []
        with open(probs) as f:
            for line in f:
                b, v = line.split("\t\t")
                bits.append(b)
                xs.append(float(v))
        case["x"] = np.array(xs, np.float64)
        case["bits"] = bits
    case["stdout"] = meta["stdout_lines"]
    return case


def read_text(case, fname):
    with open(os.path.join(case["dir"], fname)) as f:
        return f.read()


def read_bytes(case, fname):
    with open(os.path.join(case["dir"], fname), "rb") as f:
        return f.read()


def stdout_value(case, prefix):
    for ln in case["stdout"]:
        if ln.startswith(prefix):
            return ln[len(prefix):].strip()
    return None
