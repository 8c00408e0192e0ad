"""Continuous-batching request loop for box measurements on one GPU.

Each request is one periodic box (positions already in HBM) asking for
P0/P2/P4 with the line of sight along z (Pk_periodic_rsd, pyspectrum.py:460)
and B(k) (Bk_periodic, pyspectrum.py:285).  The loop

* admits a request when a worker slot is free and the device memory of all
  slots stays within the HBM budget: a busy slot counts the larger of its
  request's footprint (``footprint_bytes``) and the workspace its context
  still holds from earlier requests, an idle slot its held workspace; when a
  request does not fit, the workspace of idle slots is released first;
* runs admitted requests concurrently, one library context and one
  default-priority stream per slot, so their kernels overlap on the device;
* with ``max_batch`` > 1, batches requests continuously: a request admitted
  to a slot takes along the requests of the same shape (Ngrid, step, Ncut,
  Nmax, precision) that have arrived by then, up to ``max_batch``, and the
  slot runs them with one launch per stage (psx_bk_periodic_batch);
* computes the half field once per request and reuses it for P_l and B;
* records arrival, start and completion times (latency = completion - arrival).

Every stage sums in a fixed order, so a request's outputs are bit-identical
whether it ran alone or in the mix.
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


def footprint_bytes(Ngrid, npart, Nmax=40, step=3, Ncut=3, precision="fp32", Nmubin=10, device=0):
    """Upper estimate of the device memory one request holds: the library's
    workspace bound for its B(k) call (psx_bk_workspace_bytes), the P_l
    binning partials, and the request's own buffers (half field, outputs)."""
    ws = L.context(device).bk_workspace_bytes(Ngrid, step, Ncut, Nmax, npart, precision)
    nb = Ngrid // 2
    c = 16 if precision == "fp64" else 8
    half = (Ngrid // 2 + 1) * Ngrid * Ngrid * c
    rsd = 513 * (5 * nb + 4 * nb * Nmubin) * 8
    ntri = len(L.triangles(step, Ncut, Nmax))
    return int(ws + half + rsd + 16 * ntri * 8 + (1 << 20))


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

    def held_bytes(self):
        """Device workspace this slot keeps between requests (its context and graphs)."""
        return self.ctx.workspace_bytes() + sum(g.workspace_bytes() for g in self.graphs.values())

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
        return _result(pk, nbar, fin)

    def run_batch(self, reqs):
        """Requests of one shape as one batch: B(k) of all boxes with one launch
        per stage (psx_bk_periodic_batch), then P_l of each box from its half
        field.  Each result is bit-identical to ``run`` on that request."""
        from .batch import run_batch
        r0 = reqs[0]
        N = r0.Ngrid
        prec = L.PREC[r0.precision]
        with torch.cuda.stream(self.stream):
            out7, _, half, sums, ijl, _, keep = run_batch(
                self.ctx, [r.xyz for r in reqs], None, [r.Lbox for r in reqs], N, r0.step,
                r0.Ncut, r0.Nmax, r0.precision, self.device)
            pks = [P._pk_pbox_rsd(half[b], 2, r.Lbox, N // 2, r.Nmubin, N, prec, self.device,
                                  ctx=self.ctx) for b, r in enumerate(reqs)]
            fin = out7.cpu().numpy()[:, :, :len(ijl)]   # synchronises this stream
            del keep
        return [_result(pk, nbar, fin[b]) for b, (pk, (_, _, nbar, _)) in enumerate(zip(pks, sums))]


def _result(pk, nbar, fin):
    k, p0, p2, p4, nk = pk[:5]
    norm = (2. * np.pi) ** 3
    return {"k": k, "p0k": p0 * norm - 1. / nbar, "p2k": p2 * norm, "p4k": p4 * norm,
            "counts": nk, "b123": fin[3], "q123": fin[4], "p0k1": fin[0], "p0k2": fin[1],
            "p0k3": fin[2], "b_counts": fin[5], "b123_sn": fin[6]}


class Server:
    """Admission by slots and HBM budget; requests run concurrently on slots."""

    def __init__(self, nslots=4, hbm_budget=None, device=0, use_graphs=False, max_batch=1):
        self.device = device
        self.max_batch = max(1, min(int(max_batch), 16))
        free, total = torch.cuda.mem_get_info(device)
        self.budget = int(hbm_budget if hbm_budget is not None else 0.8 * free)
        self.use_graphs = use_graphs
        self.graph_build_s = 0.0
        self.slots = [Slot(device) for _ in range(nslots)]
        self._lock = threading.Condition()
        self._held = [0] * nslots      # workspace each slot's context holds (bytes)
        self._need = [0] * nslots      # footprint of the request a busy slot runs
        self._free_slots = list(range(nslots))
        self.peak_bytes = 0            # largest committed total at any admission

    def close(self):
        """Release the slots' graphs and library contexts."""
        for slot in self.slots:
            for g in slot.graphs.values():
                g.close()
            slot.graphs.clear()
            slot.ctx.close()
        torch.cuda.empty_cache()

    def _committed(self, s=None, need=0):
        """Device bytes of all slots, with slot ``s`` running a request of
        footprint ``need``: busy slots max(held, footprint), idle slots held."""
        tot = 0
        for t in range(len(self.slots)):
            n = need if t == s else self._need[t]
            tot += max(self._held[t], n)
        return tot

    def _admit(self, need):
        with self._lock:
            while True:
                if self._free_slots:
                    # prefer the free slot whose held workspace covers most of the request
                    s = max(self._free_slots, key=lambda t: (min(self._held[t], need), -t))
                    if self._committed(s, need) > self.budget:
                        # release idle workspace (not the chosen slot's: it is reused)
                        for t in self._free_slots:
                            if t != s and self._held[t] > 0:
                                self.slots[t].ctx.release_workspace()
                                self._held[t] = self.slots[t].held_bytes()
                    if self._committed(s, need) > self.budget and self._held[s] > 0:
                        self.slots[s].ctx.release_workspace()
                        self._held[s] = self.slots[s].held_bytes()
                    busy = any(n > 0 for n in self._need)
                    if self._committed(s, need) <= self.budget or not busy:
                        self._free_slots.remove(s)
                        self._need[s] = need
                        self.peak_bytes = max(self.peak_bytes, self._committed())
                        return s
                self._lock.wait()

    def _release(self, s, need):
        held = self.slots[s].held_bytes()
        with self._lock:
            self._held[s] = held
            self._need[s] = 0
            self._free_slots.append(s)
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
        torch.cuda.synchronize(self.device)
        with self._lock:
            self._held = [slot.held_bytes() for slot in self.slots]
        threads = []
        t0 = time.perf_counter()

        def work(group, s, need):
            # a failing request (bad input, out of memory) records its error
            # and frees its slot; the others are unaffected
            try:
                for req in group:
                    req.t_start = time.perf_counter() - t0
                if len(group) == 1:
                    results = [self.slots[s].run(group[0])]
                else:
                    results = self.slots[s].run_batch(group)
                for req, res in zip(group, results):
                    req.result = res
            except Exception as e:
                for req in group:
                    req.error = "%s: %s" % (type(e).__name__, e)
            finally:
                for req in group:
                    req.t_done = time.perf_counter() - t0
                self._release(s, need)

        def shape(r):
            return (r.Ngrid, r.step, r.Ncut, r.Nmax, r.precision)

        taken = [False] * len(reqs)
        for i, req in enumerate(reqs):
            if taken[i]:
                continue
            delay = req.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            # continuous batching: requests of the same shape that have
            # arrived by now join this one (up to max_batch)
            group, now = [i], time.perf_counter() - t0
            for j in range(i + 1, len(reqs)):
                if len(group) == self.max_batch or reqs[j].arrival > now:
                    break
                if not taken[j] and shape(reqs[j]) == shape(req):
                    group.append(j)
            for j in group:
                taken[j] = True
            grp = [reqs[j] for j in group]
            try:
                need = sum(footprint_bytes(r.Ngrid, r.xyz.shape[1], r.Nmax, r.step, r.Ncut,
                                           r.precision, r.Nmubin, self.device) for r in grp)
            except (L.PsxError, ValueError):
                need = 0  # an unsupported request fails on its slot and records the error
            s = self._admit(need)
            th = threading.Thread(target=work, args=(grp, s, need), daemon=True)
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
