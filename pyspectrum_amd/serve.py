"""Continuous-batching request loop for box measurements on one GPU.

Each request is one periodic box (positions already in HBM) asking for
P0/P2/P4 with the line of sight along z (Pk_periodic_rsd, pyspectrum.py:460)
and B(k) (Bk_periodic, pyspectrum.py:285).  The loop

* admits a request when a worker slot is free and its HBM footprint
  (``footprint_bytes``) fits the budget next to the requests in flight;
* runs admitted requests concurrently, one library context and one
  default-priority stream per slot, so their kernels overlap on the device;
* computes the half field once per request and reuses it for P_l and B;
* records arrival, start and completion times (latency = completion - arrival).

Every stage is deterministic except the float64 atomics of the P_l binning
(pk_pbox_rsd), so a request's B(k) outputs are bit-identical whether it ran
alone or in the mix, and its P_l agrees to float64 round-off.
"""
import ctypes
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import pyspectrum as P


def footprint_bytes(Ngrid, npart, Nmax=40, step=3, Ncut=3, precision="fp32"):
    """Upper estimate of the device memory a request holds (workspace + I/O)."""
    G = float(Ngrid) ** 3
    r = 8 if precision == "fp64" else 4
    ns = Nmax - Ncut // step + 1
    kx = min(Ngrid // 2 + 1, step * (Nmax + 1) + 16)
    grids = 2 * r * G + 2 * r * G / 2                      # packed FFT grid, half field
    shells = ns * G * r + 2 * 2 * r * Ngrid * Ngrid * kx * ns / 2   # D + T1/T2 (average kx)
    parts = 24.0 * npart                                   # tile keys and ranks, sorted records
    return int(grids + shells + parts + (64 << 20))


@dataclass
class Request:
    rid: int
    xyz: torch.Tensor            # float32 (3, N) on the device
    Lbox: float
    Ngrid: int
    arrival: float = 0.0         # seconds after the loop starts
    step: int = 3
    Ncut: int = 3
    Nmax: int = 40
    Nmubin: int = 10
    precision: str = "bf16x3"
    result: dict = field(default=None, repr=False)
    error: str = None            # set instead of result when the request failed
    t_start: float = float("nan")
    t_done: float = float("nan")


class Slot:
    """A worker: its own library context (workspace, plans) and stream, and
    optionally one captured B(k) graph per box shape (graphs.BkGraph)."""

    def __init__(self, device):
        self.device = device
        self.ctx = L.Context(device)
        self.graphs = {}
        with torch.cuda.device(device):
            self.stream = torch.cuda.Stream(device=device)

    def add_graph(self, key):
        from .graphs import BkGraph
        if key not in self.graphs:
            N, step, Ncut, Nmax, prec = key
            self.graphs[key] = BkGraph(N, step, Ncut, Nmax, prec, device=self.device)

    def run(self, req):
        dev = torch.device("cuda", self.device)
        prec = L.PREC[req.precision]
        N, n = req.Ngrid, req.xyz.shape[1]
        graph = self.graphs.get((N, req.step, req.Ncut, req.Nmax, req.precision))
        with torch.cuda.stream(self.stream):
            cd = torch.complex128 if req.precision == "fp64" else torch.complex64
            half = graph.half if graph is not None else torch.empty((N, N, N // 2 + 1), dtype=cd,
                                                                    device=dev)
            # each call binds this slot's stream (the current stream) to its context
            self.ctx.call("psx_fft_periodic", L.ptr(req.xyz), None, n, N, float(req.Lbox),
                          float(n), float(n), L.ptr(half), prec)
            pk = P._pk_pbox_rsd(half, 2, req.Lbox, N // 2, req.Nmubin, N, prec, self.device,
                                 ctx=self.ctx)
            ijl = L.triangles(req.step, req.Ncut, req.Nmax)
            counts = P._counts_Bk123(Ngrid=N, Nmax=req.Nmax, Ncut=req.Ncut, step=req.step,
                                     device=self.device)
            tc = P._tri_counts(counts, ijl, N)
            ntri = len(ijl)
            if graph is not None:
                graph.replay()  # shell fields, contraction, p0k on this stream
                braw, p0k = graph.braw, graph.p0k
            else:
                braw = torch.empty(max(ntri, 1), dtype=torch.float64, device=dev)
                p0k = torch.empty(req.Nmax, dtype=torch.float64, device=dev)
                self.ctx.call("psx_bk_raw", L.ptr(half), N, req.step, req.Ncut,
                       req.Nmax, L.ptr(braw), L.ptr(p0k), prec)
            nbar = n / req.Lbox ** 3
            kf = 2 * np.pi / req.Lbox
            cnt = torch.from_numpy(tc).to(dev)
            ijl_d = torch.from_numpy(np.ascontiguousarray(ijl)).to(dev)
            out7 = torch.empty((7, max(ntri, 1)), dtype=torch.float64, device=dev)
            self.ctx.call("psx_bk_finalize", L.ptr(braw), L.ptr(p0k), L.ptr(cnt),
                   L.ptr(ijl_d), ntri, N, float(nbar), (2 * np.pi) ** 3, kf ** 3, (2 * np.pi) ** 6,
                   kf ** 6, L.ptr(out7))
            fin = out7.cpu().numpy()[:, :ntri]   # synchronises this stream
        k, p0, p2, p4, nk = pk[:5]
        norm = (2. * np.pi) ** 3
        return {"k": k, "p0k": p0 * norm - 1. / nbar, "p2k": p2 * norm, "p4k": p4 * norm,
                "counts": nk, "b123": fin[3], "q123": fin[4], "p0k1": fin[0], "p0k2": fin[1],
                "p0k3": fin[2], "b_counts": fin[5], "b123_sn": fin[6]}


class Server:
    """Admission by slots and HBM budget; requests run concurrently on slots."""

    def __init__(self, nslots=4, hbm_budget=None, device=0, use_graphs=False):
        self.device = device
        free, total = torch.cuda.mem_get_info(device)
        self.budget = int(hbm_budget if hbm_budget is not None else 0.8 * free)
        self._fixed_budget = hbm_budget is not None
        self.use_graphs = use_graphs
        self.graph_build_s = 0.0
        self.slots = [Slot(device) for _ in range(nslots)]
        self._lock = threading.Condition()
        self._inflight_bytes = 0
        self._free_slots = list(range(nslots))

    def close(self):
        """Release the slots' graphs and library contexts."""
        for slot in self.slots:
            for g in slot.graphs.values():
                g.close()
            slot.graphs.clear()
            slot.ctx.close()
        torch.cuda.empty_cache()

    def _admit(self, need):
        with self._lock:
            while not self._free_slots or (self._inflight_bytes + need > self.budget
                                           and self._inflight_bytes > 0):
                self._lock.wait()
            s = self._free_slots.pop()
            self._inflight_bytes += need
            return s

    def _release(self, s, need):
        with self._lock:
            self._free_slots.append(s)
            self._inflight_bytes -= need
            self._lock.notify_all()

    def serve(self, requests):
        """Run requests at their arrival times; returns (requests, wall seconds)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        # N_tri caches are computed (or read) before the clock starts
        for key in sorted({(r.Ngrid, r.Nmax, r.Ncut, r.step) for r in reqs}):
            try:
                P._counts_Bk123(Ngrid=key[0], Nmax=key[1], Ncut=key[2], step=key[3],
                                device=self.device)
            except (L.PsxError, ValueError):
                pass  # the request fails on its own and records the error
        if self.use_graphs:
            # one captured B(k) graph per slot and box shape, built before the
            # clock; the budget then covers what is left
            tg = time.perf_counter()
            shapes = sorted({(r.Ngrid, r.step, r.Ncut, r.Nmax, r.precision) for r in reqs},
                            key=lambda k: -k[0])
            for slot in self.slots:
                for key in shapes:
                    try:
                        slot.add_graph(key)
                    except (L.PsxError, ValueError):
                        pass  # unsupported shape: its requests fail and report it
            torch.cuda.synchronize(self.device)
            self.graph_build_s += time.perf_counter() - tg
            if not self._fixed_budget:
                self.budget = int(0.8 * torch.cuda.mem_get_info(self.device)[0])
        torch.cuda.synchronize(self.device)
        threads = []
        t0 = time.perf_counter()

        def work(req, s, need):
            # a failing request (bad input, out of memory) records its error
            # and frees its slot; the others are unaffected
            try:
                req.t_start = time.perf_counter() - t0
                req.result = self.slots[s].run(req)
            except Exception as e:
                req.error = "%s: %s" % (type(e).__name__, e)
            finally:
                req.t_done = time.perf_counter() - t0
                self._release(s, need)

        for req in reqs:
            delay = req.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            need = footprint_bytes(req.Ngrid, req.xyz.shape[1], req.Nmax, req.step, req.Ncut,
                                   req.precision)
            s = self._admit(need)
            th = threading.Thread(target=work, args=(req, s, need), daemon=True)
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        return reqs, time.perf_counter() - t0


def latency_stats(reqs):
    """Latency percentiles (ms) of the requests that completed."""
    lat = np.array([r.t_done - r.arrival for r in reqs if r.error is None]) * 1e3
    if lat.size == 0:
        return {"p50_ms": float("nan"), "p99_ms": float("nan"), "mean_ms": float("nan"),
                "max_ms": float("nan")}
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


def make_stream(nreq, rate, seed, grids=(128, 192, 256, 360, 512), boxes=(1000., 2000., 2500.),
                nrange=(2e5, 4e6), redshift=0.5, sigma_v=300.0, precision="bf16x3", device=0,
                mesh=256):
    """Seeded request stream: Poisson arrivals at ``rate``/s, log-uniform N,
    lognormal positions shifted along z by Gaussian velocities (applyRSD)."""
    from . import mocks
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(math.exp(rng.uniform(math.log(nrange[0]), math.log(nrange[1]))))
        Ngrid = int(rng.choice(grids))
        Lbox = float(rng.choice(boxes))
        xyz, _ = mocks.lognormal_box(n, Lbox, seed=seed * 1000 + i, mesh=mesh, device=device)
        v = mocks.gaussian_velocities(n, sigma_v, seed=seed * 1000 + i, device=device)
        xyz = mocks.apply_rsd(xyz, v, redshift, Lbox, los=2)
        out.append(Request(i, xyz, Lbox, Ngrid, arrival=t, precision=precision))
    torch.cuda.synchronize(device)
    return out
