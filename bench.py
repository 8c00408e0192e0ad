"""Throughput / latency benchmark of the periodic-box B(k) pipeline on one MI355X.

Workload (one "step" = one box): 1,000,000 lognormal points in a 2500 Mpc/h
box (Gaussian field with P(k) ~ k^-1.5 exp(-(k/0.5)^2) on a 512^3 mesh, seed 2),
B(k) at Ngrid=360, step=3, Ncut=3, Nmax=40 (40 shells, 6350 triangles):
PCS assignment with interlacing -> 3D FFT -> fcomb -> 40 shell fields ->
triangle contraction -> N_tri normalisation, units and shot noise.  Inputs are
resident in HBM before timing; N_tri comes from the FP64 count cache (as in
the reference), computed once before the timed region.

Prints one JSON line: boxes/s of the mixed-precision (split-BF16) path, its
p50/p99 latency, the FP64 path's boxes/s and the ratio, per-stage times, the
roofline of the dominant kernel and a CPU baseline (the oracle, timed on a
bounded sample of the same workload on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm_gbs": 8000.0, "bf16_tflops": 2500.0, "fp32_tflops": 157.3, "fp64_tflops": 78.6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "fp32", "fp64"])
    ap.add_argument("--fp64-steps", type=int, default=5)
    ap.add_argument("--npart", type=int, default=1_000_000)
    ap.add_argument("--ngrid", type=int, default=360)
    ap.add_argument("--nmax", type=int, default=40)
    ap.add_argument("--step", type=int, default=3)
    ap.add_argument("--ncut", type=int, default=3)
    ap.add_argument("--lbox", type=float, default=2500.0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-triangles", type=int, default=12)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--config", default="box", choices=["box", "serve", "batch", "pmc", "survey"],
                    help="box: the single clustered box (default); serve: Poisson request "
                         "stream at --rate/s; batch: 64 uneven boxes submitted at once; "
                         "pmc: one box, a 1 GiB device copy (counter calibration), one box "
                         "-- the program to run under rocprofv3 --pmc; survey: B0_survey on a "
                         "synthetic footprint (--npart galaxies, 4x randoms)")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--nreq", type=int, default=40)
    ap.add_argument("--slots", type=int, default=4)
    return ap.parse_args()


class Box:
    """Device-resident inputs and outputs of one B(k) configuration."""

    def __init__(self, xyz, Lbox, Ngrid, step, Ncut, Nmax):
        import torch
        from pyspectrum_amd import _lib as L
        from pyspectrum_amd import pyspectrum as P
        self.L, self.torch = L, torch
        self.xyz, self.Lbox, self.N = xyz, Lbox, Ngrid
        self.step, self.Ncut, self.Nmax = step, Ncut, Nmax
        self.np_ = xyz.shape[1]
        self.ijl = L.triangles(step, Ncut, Nmax)
        self.ntri = len(self.ijl)
        counts = P._counts_Bk123(Ngrid=Ngrid, Nmax=Nmax, Ncut=Ncut, step=step)
        tc = P._tri_counts(counts, self.ijl, Ngrid)
        dev = xyz.device
        self.cnt = torch.from_numpy(tc).to(dev)
        self.ijl_d = torch.from_numpy(np.ascontiguousarray(self.ijl)).to(dev)
        self.out7 = torch.empty((7, self.ntri), dtype=torch.float64, device=dev)
        self.p0k = torch.empty(Nmax, dtype=torch.float64, device=dev)
        self.nbar = self.np_ / Lbox ** 3
        self.kf = 2 * np.pi / Lbox
        self.ctx = L.context(dev.index or 0)

    def run(self, prec):
        L = self.L
        L.call("psx_bk_periodic", self.ctx.handle, L.ptr(self.xyz), None, self.np_, self.N,
               float(self.Lbox), float(self.np_), float(self.np_), self.step, self.Ncut, self.Nmax,
               L.ptr(self.cnt), L.ptr(self.ijl_d), float(self.nbar), (2 * np.pi) ** 3, self.kf ** 3,
               (2 * np.pi) ** 6, self.kf ** 6, L.ptr(self.out7), L.ptr(self.p0k), L.PREC[prec])


def time_steps(box, prec, steps, warmup):
    torch = box.torch
    for _ in range(warmup):
        box.run(prec)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        box.run(prec)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    return wall, lat


def stage_times(box, prec, reps=3):
    """Per-stage device times (ms, median of reps) through the staged C ABI."""
    torch, L = box.torch, box.L
    N, dev = box.N, box.xyz.device
    fp = L.PREC[prec]
    cd = torch.complex128 if prec == "fp64" else torch.complex64
    rd = torch.float64 if prec == "fp64" else torch.float32
    j0 = box.Ncut // box.step
    ns = box.Nmax - j0 + 1
    half = torch.empty((N, N, N // 2 + 1), dtype=cd, device=dev)
    D = torch.empty((ns, N, N, N), dtype=rd, device=dev)
    braw = torch.empty(box.ntri, dtype=torch.float64, device=dev)
    h = box.ctx.handle
    stages = {
        "fft_periodic": lambda: L.call("psx_fft_periodic", h, L.ptr(box.xyz), None, box.np_, N,
                                       float(box.Lbox), float(box.np_), float(box.np_), L.ptr(half), fp),
        "shell_fields": lambda: L.call("psx_bk_shell_fields", h, L.ptr(half), N, box.step, box.Ncut,
                                       box.Nmax, L.ptr(D), fp),
        "contraction": lambda: L.call("psx_bk_contract", h, L.ptr(D), N, box.step, box.Ncut,
                                      box.Nmax, L.ptr(braw), fp),
    }
    out = {}
    for name, fn in stages.items():
        fn()
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(); fn(); b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        out[name] = float(np.median(ts))
    del D
    torch.cuda.empty_cache()
    return out


def accuracy(out, p0k, out64, p0k64, cnt):
    """Errors of a run against the FP64 GPU run of the same box: B before shot
    noise (B_raw = b123 + b123_sn), B with the denominator max(|B_fp64|,
    b123_sn_fp64) (B ~ 0 for shot-noise-dominated triangles), Q and P(k_j)."""
    m = cnt > 0
    braw, braw64 = out[3][m] + out[6][m], out64[3][m] + out64[6][m]
    den = np.maximum(np.abs(out64[3][m]), np.abs(out64[6][m]))
    pm = np.isfinite(p0k64) & (p0k64 != 0)
    return {"b_raw_rel_max": float(np.max(np.abs(braw - braw64) / np.abs(braw64))),
            "b_rel_max_sn_floor": float(np.max(np.abs(out[3][m] - out64[3][m]) / den)),
            "q_rel_max_sn_floor": float(np.max(np.abs(out[4][m] - out64[4][m]) /
                                               np.maximum(np.abs(out64[4][m]), 1.0))),
            "p0k_rel_max": float(np.max(np.abs(p0k[pm] - p0k64[pm]) / np.abs(p0k64[pm]))),
            "triangles": int(m.sum())}


def contraction_traffic(N, Nmax, prec):
    """HBM bytes per contraction launch measured with rocprofv3 --pmc
    (FETCH_SIZE x 2 + WRITE_SIZE, the factor 2 from a 1 GiB copy kernel), kept
    as a committed summary next to the command that produced it; None when no
    measurement exists for this workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    v = tab.get("k_contract/%s/N%d/Nmax%d" % (prec, N, Nmax))
    return None if v is None else float(v)


def cpu_baseline(xyz_np, Lbox, Ngrid, step, Ncut, Nmax, n_tri_sample):
    """The oracle (FP64 numpy restatement of the reference path, 1 thread) timed
    on a bounded sample of the same box: the full front end (assignment, FFT,
    fcomb, reflect), one shell field and n_tri_sample triangle sums; the box
    time is extrapolated to all shells and triangles."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import spectrum as S
    t0 = time.perf_counter()
    half = S.FFT_periodic(xyz_np, None, Lbox, Ngrid)
    full = S.reflect_delta(half, Ngrid)
    irk = S.shell_index(Ngrid, step)
    t_front = time.perf_counter() - t0
    t0 = time.perf_counter()
    d = [np.fft.fftn(np.where(irk == j, full, 0)).real.ravel() for j in (Nmax,)]
    t_shell = time.perf_counter() - t0
    d3 = [d[0], d[0] * 0.5, d[0] * 0.25]
    t0 = time.perf_counter()
    for _ in range(n_tri_sample):
        np.einsum("i,i,i", d3[0], d3[1], d3[2])
    t_tri = (time.perf_counter() - t0) / n_tri_sample
    ns = Nmax - Ncut // step + 1
    ntri = len(S.triangles(Nmax, Ncut, step)[0])
    box_s = t_front + ns * t_shell + ntri * t_tri
    sample_s = t_front + t_shell + n_tri_sample * t_tri
    return {"value": 1.0 / box_s, "unit": "boxes/s", "cores": 1, "kind": "port",
            "sample": ("oracle (FP64 numpy restatement) on the same 1e6-point box: full front end "
                       "%.1fs, 1 of %d shell FFTs %.1fs, %d of %d triangle einsums %.3fs each "
                       "(%.0fs sampled); box time extrapolated to %.0fs"
                       % (t_front, ns, t_shell, n_tri_sample, ntri, t_tri, sample_s, box_s))}


def cpu_baseline_all_cores(xyz_np, Lbox, Ngrid, step, Ncut, Nmax, n_tri_sample):
    """The same sampled oracle box with the host's cores (at most 16): the
    FFTs by scipy.fft (pocketfft) with one worker per core, the triangle sums
    spread over a thread pool (numpy releases the GIL in them); assignment and
    fcomb stay vectorised numpy."""
    import scipy.fft
    from concurrent.futures import ThreadPoolExecutor
    from oracle import spectrum as S
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    n = xyz_np.shape[1]
    A, B = S.assign_quad(S.clip_positions(xyz_np, Lbox), np.ones(n), Ngrid,
                         np.float32(float(Ngrid) / Lbox))
    F = scipy.fft.ifftn(A + 1j * B, workers=cores) * A.size
    F = S.fcomb_periodic(F, float(n), Ngrid)
    full = S.reflect_delta(F[: Ngrid // 2 + 1], Ngrid)
    irk = S.shell_index(Ngrid, step)
    t_front = time.perf_counter() - t0
    t0 = time.perf_counter()
    d = scipy.fft.fftn(np.where(irk == Nmax, full, 0), workers=cores).real.ravel()
    t_shell = time.perf_counter() - t0
    d3 = [d, d * 0.5, d * 0.25]
    nsum = max(n_tri_sample, 2 * cores)
    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda _: np.einsum("i,i,i", d3[0], d3[1], d3[2]), range(nsum)))
        t_tri = (time.perf_counter() - t0) / nsum
    ns = Nmax - Ncut // step + 1
    ntri = len(S.triangles(Nmax, Ncut, step)[0])
    box_s = t_front + ns * t_shell + ntri * t_tri
    return {"value": 1.0 / box_s, "unit": "boxes/s", "cores": cores, "kind": "port",
            "sample": ("oracle restatement with %d cores on the same box: front end %.1fs, 1 of %d "
                       "shell FFTs %.2fs, %d of %d triangle sums %.4fs each (threaded); box time "
                       "extrapolated to %.0fs" % (cores, t_front, ns, t_shell, nsum, ntri, t_tri,
                                                  box_s))}


def run_stream(args):
    """Serving (Poisson arrivals) or batch (all at t=0) workload through serve.Server."""
    import torch
    from pyspectrum_amd import serve
    if args.config == "serve":
        reqs = serve.make_stream(args.nreq, rate=args.rate, seed=5, precision=args.precision)
        workload = ("Poisson arrivals %.0f/s, %d redshift-space lognormal boxes, N log-uniform "
                    "[2e5, 4e6], Ngrid in {128,192,256,360,512}, Lbox in {1000,2000,2500}; "
                    "P0/P2/P4 (rsd=2, Nmubin=10) + B(k) step=3 Ncut=3 Nmax=40 per request"
                    % (args.rate, args.nreq))
    else:
        reqs = serve.make_stream(64, rate=0.0, seed=100, grids=(128, 192, 256, 360),
                                 precision=args.precision)
        workload = ("64 lognormal boxes submitted at once, N log-uniform [2e5, 4e6], Ngrid in "
                    "{128,192,256,360}, Lbox in {1000,2000,2500}; P0/P2/P4 + B(k) Nmax=40")
    def measure(use_graphs):
        server = serve.Server(nslots=args.slots, use_graphs=use_graphs)
        # warm the plans of every Ngrid on every slot
        warm = [serve.Request(-1, r.xyz, r.Lbox, r.Ngrid, precision=args.precision)
                for r in {r.Ngrid: r for r in reqs}.values()]
        for s in server.slots:
            for w in warm:
                s.run(w)
        torch.cuda.synchronize()
        out, wall = server.serve(reqs)
        failed = [r for r in out if r.error is not None]
        if failed:
            raise SystemExit("request %d failed: %s" % (failed[0].rid, failed[0].error))
        res = (len(out) / wall, serve.latency_stats(out), wall, server.graph_build_s)
        server.close()
        return res

    extra = {}
    if args.config == "batch":
        # eager slots, then the same batch with one captured B(k) graph per
        # slot and Ngrid (built before the clock)
        rate, st, wall, _ = measure(False)
        rate_g, st_g, wall_g, tg = measure(True)
        extra = {"graph_boxes_per_s": round(rate_g, 3), "graph_p50_ms": round(st_g["p50_ms"], 2),
                 "graph_p99_ms": round(st_g["p99_ms"], 2), "graph_build_s": round(tg, 3)}
    else:
        rate, st, wall, _ = measure(False)
    line = {
        "metric": "requests/s and p50/p99 latency" if args.config == "serve" else "boxes/s",
        "value": round(rate, 3), "unit": "boxes/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": args.precision, "data": "synthetic lognormal mocks",
        "config": {"workload": workload, "slots": args.slots},
        "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
        "mean_ms": round(st["mean_ms"], 2), "wall_s": round(wall, 3)}
    line.update(extra)
    print(json.dumps(line))


def run_survey(args):
    """Survey-geometry B(k) (B0_survey): a synthetic footprint (RA 0-90 deg,
    Dec -30..30 deg, z 0.2-0.6, n(z) smooth) of --npart galaxies and 4x as many
    randoms at Ngrid=360, Lbox=3600 Mpc/h.  Times the whole call (catalogues
    copied to the device, positions/weights/sums by psx_survey_catalogue,
    assignment, FFT, shells, contraction) for the split-BF16 and FP64
    pipelines, and for comparison the reference-style numpy catalogue
    preparation of both catalogues on the host."""
    import torch
    from pyspectrum_amd import pyspectrum as P
    from pyspectrum_amd import util
    rng = np.random.default_rng(args.seed)
    cosmo = util.default_cosmology()

    def catalogue(n):
        u = rng.uniform(size=n)
        dec = np.degrees(np.arcsin(np.sin(np.radians(-30.)) + u * 2 * np.sin(np.radians(30.))))
        radecz = np.array([rng.uniform(0, 90, n), dec, rng.uniform(0.2, 0.6, n)])
        nb = 3e-4 * np.exp(-((radecz[2] - 0.4) / 0.15) ** 2)
        return radecz, nb

    nd = args.npart
    d, nbd = catalogue(nd)
    r, nbr = catalogue(4 * nd)
    kw = dict(P0_fkp=1e4, Lbox=3600., Ngrid=360, step=args.step, Ncut=args.ncut, Nmax=args.nmax,
              cosmo=cosmo)
    # CPU baseline leg: the oracle's host restatement of the reference's
    # catalogue preparation (pyspectrum.py:767-806) for both catalogues
    from oracle import spectrum as S
    t0 = time.perf_counter()
    S.survey_catalogue(d, nbd, None, 1e4, 3600., cosmo)
    S.survey_catalogue(r, nbr, None, 1e4, 3600., cosmo)
    host_ms = (time.perf_counter() - t0) * 1e3
    res, times = {}, {}
    for prec in (args.precision, "fp64"):
        P.B0_survey(d, nbd, radecz_r=r, nbar_r=nbr, precision=prec, **kw)  # warm-up (plans, counts)
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            res[prec] = P.B0_survey(d, nbd, radecz_r=r, nbar_r=nbr, precision=prec, **kw)
            ts.append((time.perf_counter() - t0) * 1e3)
        times[prec] = float(np.median(ts))
    a, b = res[args.precision], res["fp64"]
    big = np.max(np.abs(b["b123"]))
    err = float(np.max(np.abs(a["b123"] - b["b123"])) / big)
    print(json.dumps({
        "metric": "survey B(k) catalogues/s (B0_survey, data + 4x randoms)",
        "value": round(1e3 / times[args.precision], 3), "unit": "catalogues/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": args.precision, "data": "synthetic survey footprint",
        "config": {"workload": "%d galaxies + %d randoms, RA 0-90, Dec -30..30, z 0.2-0.6, "
                               "Ngrid=360, Lbox=3600, step=%d, Ncut=%d, Nmax=%d"
                               % (nd, 4 * nd, args.step, args.ncut, args.nmax)},
        "ms_per_call": round(times[args.precision], 2), "fp64_ms_per_call": round(times["fp64"], 2),
        "host_numpy_catalogue_ms": round(host_ms, 2),
        "b123_err_vs_fp64_rel_to_max": err, "triangles": int(len(a["b123"]))}))


def run_pmc(args):
    """Short program for hardware-counter passes: the box once (plans, counts
    and buffers warm), a 1 GiB device-to-device copy whose FETCH_SIZE and
    WRITE_SIZE calibrate the counters, then the box once more."""
    import torch
    from pyspectrum_amd import mocks
    xyz, _ = mocks.lognormal_box(args.npart, args.lbox, seed=args.seed)
    box = Box(xyz, args.lbox, args.ngrid, args.step, args.ncut, args.nmax)
    box.run(args.precision)
    torch.cuda.synchronize()
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    b = a.clone()
    torch.cuda.synchronize()
    del a, b
    box.run(args.precision)
    torch.cuda.synchronize()
    print("pmc program done: %s, Ngrid=%d, Nmax=%d" % (args.precision, args.ngrid, args.nmax))


def main():
    args = parse()
    if args.gpus == 1 and args.config == "pmc":
        return run_pmc(args)
    if args.gpus == 1 and args.config == "survey":
        return run_survey(args)
    if args.gpus == 1 and args.config != "box":
        return run_stream(args)
    if args.gpus != 1:
        print("bench.py measures one GPU: every box runs on a single MI355X; "
              "--gpus %d is not supported (multi-GPU sharding is design-only)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    from pyspectrum_amd import _lib as L
    from pyspectrum_amd import mocks
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    Lbox, N, step, Ncut, Nmax = args.lbox, args.ngrid, args.step, args.ncut, args.nmax
    t0 = time.perf_counter()
    xyz, A = mocks.lognormal_box(args.npart, Lbox, seed=args.seed)
    torch.cuda.synchronize()
    if args.verbose:
        print("mock generated in %.1fs (A=%.4g)" % (time.perf_counter() - t0, A), file=sys.stderr, flush=True)
    box = Box(xyz, Lbox, N, step, Ncut, Nmax)
    if args.verbose:
        print("counts ready", file=sys.stderr, flush=True)

    wall, lat = time_steps(box, args.precision, args.steps, args.warmup)
    value = args.steps / wall
    out_mixed, p0k_mixed = box.out7.cpu().numpy().copy(), box.p0k.cpu().numpy().copy()
    wall64, lat64 = time_steps(box, "fp64", args.fp64_steps, 1)
    v64 = args.fp64_steps / wall64
    err = accuracy(out_mixed, p0k_mixed, box.out7.cpu().numpy(), box.p0k.cpu().numpy(),
                   box.cnt.cpu().numpy())
    st = stage_times(box, args.precision)
    st64 = stage_times(box, "fp64")
    if args.verbose:
        print("stages", st, st64, file=sys.stderr, flush=True)

    # roofline of the dominant kernel (the triangle contraction)
    G = float(N) ** 3
    ijl = box.ijl
    npairs = len({(int(i), int(j)) for i, j, _ in ijl})
    flops = G * (npairs + 2 * box.ntri)
    t_c = st["contraction"] / 1e3
    achieved = flops / t_c / 1e12
    peak = PEAK["bf16_tflops"] if args.precision == "bf16x3" else (
        PEAK["fp32_tflops"] if args.precision == "fp32" else PEAK["fp64_tflops"])
    roofline = {"bound": "mfma", "kernel": "k_contract", "achieved": round(achieved, 2),
                "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                "traffic": contraction_traffic(N, Nmax, args.precision)}

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        xyz_np = xyz.double().cpu().numpy()
        cpu = cpu_baseline(xyz_np, Lbox, N, step, Ncut, Nmax, args.cpu_sample_triangles)
        cpu_all = cpu_baseline_all_cores(xyz_np, Lbox, N, step, Ncut, Nmax,
                                         args.cpu_sample_triangles)
    out = {
        "metric": "B(k) boxes/s at Ngrid=%d + p50/p99 latency, mixed-precision vs FP64 GPU" % N,
        "value": round(value, 3), "unit": "boxes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3+f32", "fp32": "f32", "fp64": "f64"}[args.precision],
        "data": "synthetic lognormal mock (seed %d), generated on the GPU" % args.seed,
        "config": {"workload": "single clustered box: %.3g lognormal points, Lbox=%g, Ngrid=%d, "
                               "step=%d, Ncut=%d, Nmax=%d (%d triangles)"
                               % (args.npart, Lbox, N, step, Ncut, Nmax, box.ntri),
                   "precision": args.precision},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "fp64_boxes_per_s": round(v64, 3),
        "fp64_p50_ms": round(float(np.percentile(lat64, 50)), 3),
        "mixed_vs_fp64": round(value / v64, 3),
        "error_vs_fp64": err,
        "stages_ms": {k: round(v, 3) for k, v in st.items()},
        "stages_fp64_ms": {k: round(v, 3) for k, v in st64.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
