"""Build-time autotuner for the compile-time tile shapes and FFT radix order.

The kernels are specialised at compile time -- waves per workgroup and pair
tiles per wave of the split-BF16 contractions (fused and staged), rows per
float32 accumulation chain, row pairs per tile of the shell x pass, lines per
FFT tile, the radix order -- so a tuning choice is a set of preprocessor
definitions.  Two steps:

    python -m pyspectrum_amd.tune build            # CPU: one library per candidate
    python -m pyspectrum_amd.tune time [--ngrid 360 --nmax 40 ...]   # GPU

``build`` compiles every candidate of ``CANDIDATES`` into
``pyspectrum_amd/_tune/<name>/libpsx.so``.  ``time`` times the B(k) stages of
each candidate (FFT_periodic, then shell fields + contraction + shell powers
as psx_bk_raw runs them; device events,
median of repeats) on a seeded lognormal box, each in its own process (a
process loads one library), over several interleaved rounds, and writes the
fastest candidate's definitions -- the default unless another is faster by
a clear margin -- with every candidate's times, to
``pyspectrum_amd/tuning.json``; ``build()`` applies them from then on.  Every candidate computes the same estimator
(only tile shapes and orders change), within the stated precision.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNE_DIR = os.path.join(HERE, "_tune")

# name -> definitions; the empty set is the built-in default
CANDIDATES = {
    "default": {},
    # split-BF16 contraction fused with the shell x pass (psx_xcontract.hip)
    "xc_flush2": {"PSX_XC_FLUSH": 2},
    "xc_flush4": {"PSX_XC_FLUSH": 4},
    "xc_w8x3": {"PSX_XC_WAVES": 8, "PSX_XC_MAXPT": 3},
    # FFT tiles and radix order (front end, shell z/y passes, fused x pass)
    "fft_lines8": {"PSX_TILE_LINES_F32": 8},
    "fft_order_5_9_8": {"PSX_FFT_ORDER": 0},
    # the staged contraction (grids or shell counts outside the fused kernel)
    "ct_w8x3": {"PSX_CT_WAVES_BF": 8, "PSX_CT_MAXPT_BF": 3},
    "shellx_p8": {"PSX_SHELL_X_PAIRS": 8},
}


def build_candidates(names=None, verbose=True):
    from . import build as B
    for name in names or CANDIDATES:
        d = CANDIDATES[name]
        out = os.path.join(TUNE_DIR, name)
        os.makedirs(out, exist_ok=True)
        t0 = time.time()
        objdir = os.path.join(out, "obj")
        B.build(defines=["-D%s=%d" % kv for kv in sorted(d.items())],
                lib=os.path.join(out, "libpsx.so"), objdir=objdir)
        shutil.rmtree(objdir, ignore_errors=True)
        if verbose:
            print("built %s %s (%.0fs)" % (name, d, time.time() - t0), flush=True)


def _time_one(lib, ngrid, npart, lbox, step, ncut, nmax, prec, reps, seed):
    """Stage times (ms) of one library; runs in a child process."""
    from . import _lib as L
    L.LIB_PATH = lib
    import numpy as np
    import torch
    from . import mocks
    dev = torch.device("cuda", 0)
    xyz, _ = mocks.lognormal_box(npart, lbox, seed=seed)
    ctx = L.Context(0)
    fp = L.PREC[prec]
    cd = torch.complex128 if prec == "fp64" else torch.complex64
    half = torch.empty((ngrid, ngrid, ngrid // 2 + 1), dtype=cd, device=dev)
    braw = torch.empty(max(len(L.triangles(step, ncut, nmax)), 1), dtype=torch.float64, device=dev)
    p0k = torch.empty(nmax, dtype=torch.float64, device=dev)
    n = xyz.shape[1]
    # the two halves of a box: the front end, and shell fields + contraction +
    # shell powers (fused x pass and contraction where the shape allows)
    stages = {
        "fft_periodic": lambda h: L.call("psx_fft_periodic", h, L.ptr(xyz), None, n, ngrid,
                                         float(lbox), float(n), float(n), L.ptr(half), fp),
        "bk_raw": lambda h: L.call("psx_bk_raw", h, L.ptr(half), ngrid, step, ncut, nmax,
                                   L.ptr(braw), L.ptr(p0k), fp),
    }
    out = {}
    for name, fn in stages.items():
        fn(ctx.handle)  # warm-up (plans, workspace)
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn(ctx.handle)
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        out[name] = float(np.median(ts))
    out["total"] = sum(out.values())
    return out


def time_candidates(args):
    """Time every built candidate in ``args.rounds`` interleaved rounds (one
    child process per candidate and round); a candidate replaces the default
    only if its median total beats the default's by ``args.min_gain``
    (process-to-process spread of identical code is ~2 %)."""
    runs = {}
    for rnd in range(args.rounds):
        for name in CANDIDATES:
            lib = os.path.join(TUNE_DIR, name, "libpsx.so")
            if not os.path.isfile(lib):
                if rnd == 0:
                    print("skip %s: not built (run `python -m pyspectrum_amd.tune build`)" % name)
                continue
            cmd = [sys.executable, "-m", "pyspectrum_amd.tune", "_child", lib, str(args.ngrid),
                   str(args.npart), str(args.lbox), str(args.step), str(args.ncut), str(args.nmax),
                   args.precision, str(args.reps), str(args.seed)]
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=args.timeout,
                               cwd=os.path.dirname(HERE))
            if r.returncode != 0:
                print("candidate %s failed:\n%s" % (name, r.stderr[-2000:]), flush=True)
                continue
            res = json.loads(r.stdout.strip().splitlines()[-1])
            runs.setdefault(name, []).append(res)
            print("round %d %-18s %s" % (rnd, name, {k: round(v, 3) for k, v in res.items()}),
                  flush=True)
    if not runs:
        raise SystemExit("no candidate timed")
    med = {n: float(sorted(r["total"] for r in rs)[len(rs) // 2]) for n, rs in runs.items()}
    best = min(med, key=med.get)
    if "default" in med and med[best] > med["default"] * (1.0 - args.min_gain):
        best = "default"
    doc = {"defines": CANDIDATES[best], "candidate": best, "median_total_ms": med,
           "min_gain": args.min_gain,
           "workload": {"ngrid": args.ngrid, "npart": args.npart, "lbox": args.lbox,
                        "step": args.step, "ncut": args.ncut, "nmax": args.nmax,
                        "precision": args.precision, "seed": args.seed},
           "stage_ms": runs}
    out = args.out or os.path.join(HERE, "tuning.json")
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    print("best: %s %s (median %.3f ms) -> %s" % (best, CANDIDATES[best], med[best], out), flush=True)
    return doc


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "_child":
        lib, ngrid, npart, lbox, step, ncut, nmax, prec, reps, seed = argv[1:11]
        res = _time_one(lib, int(ngrid), int(float(npart)), float(lbox), int(step), int(ncut),
                        int(nmax), prec, int(reps), int(seed))
        print(json.dumps(res))
        return
    ap = argparse.ArgumentParser(prog="python -m pyspectrum_amd.tune")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build", help="compile every candidate (CPU)")
    b.add_argument("names", nargs="*")
    t = sub.add_parser("time", help="time every built candidate (GPU), write tuning.json")
    t.add_argument("--ngrid", type=int, default=360)
    t.add_argument("--npart", type=float, default=1e6)
    t.add_argument("--lbox", type=float, default=2500.0)
    t.add_argument("--step", type=int, default=3)
    t.add_argument("--ncut", type=int, default=3)
    t.add_argument("--nmax", type=int, default=40)
    t.add_argument("--precision", default="bf16x3")
    t.add_argument("--reps", type=int, default=5)
    t.add_argument("--seed", type=int, default=2)
    t.add_argument("--timeout", type=int, default=300)
    t.add_argument("--rounds", type=int, default=3)
    t.add_argument("--min-gain", type=float, default=0.02)
    t.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "build":
        build_candidates(a.names or None)
    else:
        time_candidates(a)


if __name__ == "__main__":
    main()
