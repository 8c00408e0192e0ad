"""Throughput / latency benchmark of the periodic-box B(k) pipeline on one MI355X.

Workload (one "step" = one box): 1,000,000 lognormal points in a 2500 Mpc/h
box (Gaussian field with P(k) ~ k^-1.5 exp(-(k/0.5)^2) on a 512^3 mesh, seed 2),
B(k) at Ngrid=360, step=3, Ncut=3, Nmax=40 (40 shells, 6350 triangles):
PCS assignment with interlacing -> 3D FFT -> fcomb -> 40 shell fields ->
triangle contraction -> N_tri normalisation, units and shot noise.  Inputs are
resident in HBM before timing; N_tri comes from the FP64 count cache (as in
the reference), computed once before the timed region.

Prints one JSON line: boxes/s of the mixed-precision (split-BF16) path, its
p50/p99 latency, the FP64 path's boxes/s and the ratio, the device time per
box of each main kernel (HIP event pairs the library's dispatches record over
the timed region), the roofline of the kernel with the most time and a
CPU baseline (the reference-structured oracle path, timed on a bounded
sample of the same box on this host, 1 thread and all cores).
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
    ap.add_argument("--config", default="box", choices=["box", "serve", "batch", "pmc", "survey", "cpu"],
                    help="box: the single clustered box (default); serve: Poisson request "
                         "stream at --rate/s; batch: 64 uneven boxes submitted at once; "
                         "pmc: one box, a 1 GiB device copy (counter calibration), one box "
                         "-- the program to run under rocprofv3 --pmc; survey: B0_survey on a "
                         "synthetic footprint (--npart galaxies, 4x randoms); cpu: full host "
                         "timings of the reference-structured CPU path (plain config at 1 "
                         "thread and all cores, the clustered box at all cores)")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--nreq", type=int, default=40)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=16,
                    help="serve/batch: same-shape requests batched per slot (1 = no batching)")
    ap.add_argument("--batched-only", action="store_true",
                    help="batch: only the batched measurement (for kernel traces)")
    ap.add_argument("--batch-slots", type=int, default=2,
                    help="batch: slots of the batched measurement (a 16-box batch at Ngrid 360 "
                         "holds about 100 GB of workspace)")
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


# Kernels whose launches carry HIP start/stop events while kernel timing is on.
KERNELS = ("k_assign_lines", "k_fft_lines", "k_fft_zcomb", "k_shell_z", "k_shell_y", "k_shell_x",
           "k_contract", "k_xcontract")


def time_steps(box, prec, steps, warmup):
    """Wall time and per-step latency of `steps` boxes, and the device time
    of each main kernel over the same timed region: each launch carries a
    HIP event pair that its dispatch records, on the stream the kernels run
    on (the context's stream, bound to torch's current stream by each
    call)."""
    torch = box.torch
    for _ in range(warmup):
        box.run(prec)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    box.ctx.kernel_timing(True)
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        box.run(prec)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = {}
    for k in KERNELS:
        n, ms = box.ctx.kernel_time(k)
        if n:
            kern[k] = {"launches_per_box": n / steps, "ms_per_launch": ms / n,
                       "ms_per_box": ms / steps}
    box.ctx.kernel_timing(False)
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    return wall, lat, kern


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


def kernel_traffic(kernel, N, Nmax, prec):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a 1 GiB copy),
    kept as a committed summary next to the command that produced it; None
    when no measurement exists for this kernel and workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    v = tab.get("%s/%s/N%d/Nmax%d" % (kernel, prec, N, Nmax))
    return None if v is None else float(v)


def roofline(box, prec, kern):
    """Roofline of the kernel with the most device time per box.  The
    triangle sums are G * (Npairs + 2 Ntri) flops per launch (one product per
    shell pair and point, one multiply-add per triangle and point); the fused
    kernel also runs the x pass of the shell fields, 5 N log2(N) flops per
    complex line of N points, ceil(Nshell / 2) lines per (y, z) row."""
    name = max(kern, key=lambda k: kern[k]["ms_per_box"])
    N = box.N
    G = float(N) ** 3
    npairs = len({(int(i), int(j)) for i, j, _ in box.ijl})
    flops = G * (npairs + 2 * box.ntri)
    if name == "k_xcontract":
        ns = box.Nmax - box.Ncut // box.step + 1
        flops += 5.0 * G * np.log2(N) * ((ns + 1) // 2)
    elif name != "k_contract":
        return {"bound": None, "kernel": name, "achieved": None, "peak": None, "unit": None,
                "frac": None, "traffic": kernel_traffic(name, N, box.Nmax, prec)}
    t = kern[name]["ms_per_launch"] / 1e3
    achieved = flops / t / 1e12
    peak = PEAK["bf16_tflops"] if prec == "bf16x3" else (
        PEAK["fp32_tflops"] if prec == "fp32" else PEAK["fp64_tflops"])
    return {"bound": "mfma", "kernel": name, "achieved": round(achieved, 2), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "flops_per_launch": flops, "ms_per_launch": round(kern[name]["ms_per_launch"], 4),
            "traffic": kernel_traffic(name, N, box.Nmax, prec)}


def host_cores():
    """Cores this process may use on the host, at most 16 (one GPU's share)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_box(xyz_np, Lbox, Ngrid, step, Ncut, Nmax, cores, ntri_sample=None, say=None):
    """The reference-structured CPU path (the oracle's restatement) on one box
    with `cores` threads: float32 PCS + interlaced assignment, one complex64
    packed c2c backward FFT (scipy.fft, `cores` workers), fcomb, the
    Hermitian reflection, one complex64 forward FFT per shell stored as
    float64, and one float64 triple sum per triangle (a thread pool over
    triangles; numpy releases the GIL in them).  With ntri_sample, only the
    first shell FFT and that many triangle sums run and the shell and
    triangle totals are extrapolated.  Returns the stage times (s)."""
    import scipy.fft
    from concurrent.futures import ThreadPoolExecutor
    from oracle import spectrum as S
    t = {}
    t0 = time.perf_counter()
    n = xyz_np.shape[1]
    A, B = S.assign_quad(S.clip_positions(xyz_np, Lbox), np.ones(n), Ngrid,
                         np.float32(float(Ngrid) / Lbox))
    F = scipy.fft.ifftn((A + 1j * B).astype(np.complex64), workers=cores) * np.float32(A.size)
    del A, B
    F = S.fcomb_periodic(F.astype(np.complex128), float(n), Ngrid)
    full = S.reflect_delta(F[: Ngrid // 2 + 1], Ngrid)
    del F
    irk = S.shell_index(Ngrid, step)
    t["front"] = time.perf_counter() - t0
    j0 = Ncut // step
    shells = list(range(j0, Nmax + 1))
    run_shells = shells[-1:] if ntri_sample else shells
    t0 = time.perf_counter()
    D = {}
    for j in run_shells:
        D[j] = scipy.fft.fftn(np.where(irk == j, full, 0).astype(np.complex64),
                              workers=cores).real.astype(np.float64).ravel()
        if say:
            say("  shell %d done (%.1fs)" % (j, time.perf_counter() - t0))
    t["shells"] = (time.perf_counter() - t0) * len(shells) / len(run_shells)
    I, J, L, _ = S.triangles(Nmax, Ncut, step)
    ntri = len(I)
    if ntri_sample:
        d = D[shells[-1]]
        D = {j: d for j in shells}
    todo = range(min(ntri, ntri_sample)) if ntri_sample else range(ntri)

    def one(k):
        return np.einsum("i,i,i", D[I[k]], D[J[k]], D[L[k]])

    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        for c, _ in enumerate(ex.map(one, todo)):
            if say and c % 1000 == 999:
                say("  %d triangles (%.1fs)" % (c + 1, time.perf_counter() - t0))
    t["triangles"] = (time.perf_counter() - t0) * ntri / len(todo)
    t["total"] = t["front"] + t["shells"] + t["triangles"]
    t["ntri"] = ntri
    t["nshells"] = len(shells)
    return t


def cpu_baseline(xyz_np, Lbox, Ngrid, step, Ncut, Nmax, n_tri_sample, cores):
    """Bounded sample of the box on the host (about 10-30 s): the full front
    end, 1 of the shell FFTs and n_tri_sample triangle sums, extrapolated."""
    t = cpu_box(xyz_np, Lbox, Ngrid, step, Ncut, Nmax, cores, ntri_sample=n_tri_sample)
    return {"value": 1.0 / t["total"], "unit": "boxes/s", "cores": cores, "kind": "port",
            "sample": ("reference-structured oracle path on the same box with %d thread(s) (%s): "
                       "full front end %.1fs, 1 of %d complex64 shell FFTs (x%d = %.1fs), %d of "
                       "%d float64 triangle sums (extrapolated %.1fs); box time %.1fs"
                       % (cores, cpu_model(), t["front"], t["nshells"], t["nshells"], t["shells"],
                          n_tri_sample, t["ntri"], t["triangles"], t["total"]))}


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
    def measure(use_graphs, max_batch=1, nslots=args.slots):
        server = serve.Server(nslots=nslots, use_graphs=use_graphs, max_batch=max_batch)
        # warm the plans of every Ngrid on every slot
        warm = [serve.Request(-1, r.xyz, r.Lbox, r.Ngrid, precision=args.precision)
                for r in {r.Ngrid: r for r in reqs}.values()]
        for s in server.slots:
            for w in warm:
                s.run(w)
        torch.cuda.synchronize()
        if args.config == "batch":
            # an untimed pass over the same requests: the slots' workspaces
            # reach the size their batches need before the clock starts
            server.serve([serve.Request(r.rid, r.xyz, r.Lbox, r.Ngrid, arrival=r.arrival,
                                        precision=r.precision) for r in reqs])
            torch.cuda.synchronize()
        # batch config: the median of three timed passes
        runs = []
        for _ in range(3 if args.config == "batch" else 1):
            rr = [serve.Request(r.rid, r.xyz, r.Lbox, r.Ngrid, arrival=r.arrival,
                                precision=r.precision) for r in reqs]
            out, wall = server.serve(rr)
            failed = [r for r in out if r.error is not None]
            if failed:
                raise SystemExit("request %d failed: %s" % (failed[0].rid, failed[0].error))
            runs.append((wall, out))
        wall, out = sorted(runs, key=lambda x: x[0])[len(runs) // 2]
        res = (len(out) / wall, serve.latency_stats(out), wall, server.graph_build_s)
        server.close()
        return res

    extra = {}
    if args.config == "batch":
        # the headline: same-shape requests batched (one launch per stage per
        # batch, up to --max-batch boxes); then one box per slot, eager and
        # with one captured B(k) graph per slot and Ngrid (built before the clock)
        rate, st, wall, _ = measure(False, args.max_batch, args.batch_slots)
        extra = {"max_batch": args.max_batch, "batch_slots": args.batch_slots}
        if not args.batched_only:
            rate_1, st_1, _, _ = measure(False)
            rate_g, st_g, wall_g, tg = measure(True)
            extra.update({"unbatched_slots": args.slots,
                          "unbatched_boxes_per_s": round(rate_1, 3),
                          "unbatched_p50_ms": round(st_1["p50_ms"], 2),
                          "unbatched_p99_ms": round(st_1["p99_ms"], 2),
                          "graph_boxes_per_s": round(rate_g, 3),
                          "graph_p50_ms": round(st_g["p50_ms"], 2),
                          "graph_p99_ms": round(st_g["p99_ms"], 2), "graph_build_s": round(tg, 3)})
    else:
        rate, st, wall, _ = measure(False, args.max_batch)
        rate_1, st_1, _, _ = measure(False)
        extra = {"max_batch": args.max_batch, "unbatched_boxes_per_s": round(rate_1, 3),
                 "unbatched_p50_ms": round(st_1["p50_ms"], 2),
                 "unbatched_p99_ms": round(st_1["p99_ms"], 2)}
    line = {
        "metric": "requests/s and p50/p99 latency" if args.config == "serve" else "boxes/s",
        "value": round(rate, 3), "unit": "boxes/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": args.precision, "data": "synthetic lognormal mocks",
        "config": {"workload": workload,
                   "slots": args.batch_slots if args.config == "batch" else args.slots},
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


def run_cpu(args):
    """Full host timings of the reference-structured CPU path, nothing
    sampled: the plain configuration (1e5 uniform points, seed 1, Lbox 1000,
    Ngrid 96, step 3, Ncut 3, Nmax 30; P(k) and B(k)) with 1 thread and with
    all cores, and the single clustered box (1e6 lognormal points, seed 2,
    Lbox 2500, Ngrid 360, Nmax 40) with all cores; the GPU's FP64 and
    split-BF16 B(k) of the plain configuration through the Python API beside
    them.  Progress goes to stderr, one JSON line to stdout."""
    import torch
    from oracle import spectrum as S
    from pyspectrum_amd import mocks
    from pyspectrum_amd import pyspectrum as P

    def say(m):
        print(m, file=sys.stderr, flush=True)

    cores = host_cores()
    res = {"metric": "reference-structured CPU path, full boxes", "cpu_model": cpu_model(),
           "cores_all": cores}
    rng = np.random.default_rng(1)
    xyz1 = rng.uniform(0, 1000.0, size=(3, 100000)).astype(np.float32)
    for c in (1, cores):
        say("plain config, %d thread(s)" % c)
        t0 = time.perf_counter()
        S.Pk_periodic(xyz1.astype(np.float64), None, 1000.0, 96)
        tpk = time.perf_counter() - t0
        t = cpu_box(xyz1.astype(np.float64), 1000.0, 96, 3, 3, 30, c)
        res["plain_%d_threads" % c] = {"pk_s": round(tpk, 3), "bk_s": round(t["total"], 3),
                                       "front_s": round(t["front"], 3),
                                       "shells_s": round(t["shells"], 3),
                                       "triangles_s": round(t["triangles"], 3)}
    for prec in ("fp64", "bf16x3"):
        P.Bk_periodic(xyz1, Lbox=1000.0, Ngrid=96, step=3, Ncut=3, Nmax=30, precision=prec)
        torch.cuda.synchronize()
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            P.Bk_periodic(xyz1, Lbox=1000.0, Ngrid=96, step=3, Ncut=3, Nmax=30, precision=prec)
            ts.append(time.perf_counter() - t0)
        res["plain_gpu_%s_call_s" % prec] = round(float(np.median(ts)), 5)
    say("clustered box, %d threads" % cores)
    xyz, _ = mocks.lognormal_box(1_000_000, 2500.0, seed=2)
    xyz2 = xyz.double().cpu().numpy()
    del xyz
    t = cpu_box(xyz2, 2500.0, 360, 3, 3, 40, cores, say=say)
    res["clustered_%d_threads" % cores] = {
        "bk_s": round(t["total"], 2), "front_s": round(t["front"], 2),
        "shells_s": round(t["shells"], 2), "triangles_s": round(t["triangles"], 2),
        "boxes_per_s": 1.0 / t["total"]}
    print(json.dumps(res))


def main():
    args = parse()
    if args.gpus == 1 and args.config == "pmc":
        return run_pmc(args)
    if args.gpus == 1 and args.config == "survey":
        return run_survey(args)
    if args.gpus == 1 and args.config == "cpu":
        return run_cpu(args)
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

    wall, lat, kern = time_steps(box, args.precision, args.steps, args.warmup)
    value = args.steps / wall
    out_mixed, p0k_mixed = box.out7.cpu().numpy().copy(), box.p0k.cpu().numpy().copy()
    wall64, lat64, kern64 = time_steps(box, "fp64", args.fp64_steps, 1)
    v64 = args.fp64_steps / wall64
    err = accuracy(out_mixed, p0k_mixed, box.out7.cpu().numpy(), box.p0k.cpu().numpy(),
                   box.cnt.cpu().numpy())
    roof = roofline(box, args.precision, kern)
    if args.verbose:
        print("kernels", kern, kern64, file=sys.stderr, flush=True)

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        xyz_np = xyz.double().cpu().numpy()
        cpu = cpu_baseline(xyz_np, Lbox, N, step, Ncut, Nmax, args.cpu_sample_triangles, 1)
        cpu_all = cpu_baseline(xyz_np, Lbox, N, step, Ncut, Nmax, 4 * args.cpu_sample_triangles,
                               host_cores())
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
        "kernels_ms_per_box": {k: round(v["ms_per_box"], 4) for k, v in kern.items()},
        "kernels_fp64_ms_per_box": {k: round(v["ms_per_box"], 4) for k, v in kern64.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
