"""HIP-graph replay of the B(k) stages for repeated boxes of one shape.

The stages after FFT_periodic -- the shell fields (pruned inverse FFTs of each
shell), the triangle contraction with its fixed-order reduction, and the shell
powers p0k -- depend only on (Ngrid, step, Ncut, Nmax, precision) and on buffer
addresses, not on the particles.  ``BkGraph`` captures that sequence once in a
HIP graph (torch.cuda.CUDAGraph over the library's launches) and replays it for
every box of that shape: one graph launch instead of about a dozen kernel
launches and their host work (pyspectrum.py:384-457 is the loop it replaces).

Each graph owns a library context, so its workspace (shell fields, FFT
intermediates, contraction partials) is allocated once before capture and is
never reallocated under it; the half field it reads (``half``) and its
outputs (``braw``, ``p0k``) are fixed tensors.  The sequence holds no memset
nodes (counters and accumulators are cleared by fill kernels) and no host
synchronisation.
"""
import torch

from . import _lib as L


class BkGraph:
    """Captured shell-field + contraction + p0k sequence for one box shape."""

    def __init__(self, Ngrid, step=3, Ncut=3, Nmax=40, precision="bf16x3", device=0):
        self.key = (Ngrid, step, Ncut, Nmax, precision)
        self.N, self.step, self.Ncut, self.Nmax = Ngrid, step, Ncut, Nmax
        self.precision = precision
        self.prec = L.PREC[precision]
        self.device = device
        dev = torch.device("cuda", device)
        self.ctx = L.Context(device)
        cd = torch.complex128 if precision == "fp64" else torch.complex64
        self.ntri = len(L.triangles(step, Ncut, Nmax))
        self.half = torch.zeros((Ngrid, Ngrid, Ngrid // 2 + 1), dtype=cd, device=dev)
        self.braw = torch.zeros(max(self.ntri, 1), dtype=torch.float64, device=dev)
        self.p0k = torch.zeros(Nmax, dtype=torch.float64, device=dev)
        self.graph = torch.cuda.CUDAGraph()
        self.stream = torch.cuda.Stream(device=dev)
        self.stream.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(self.stream):
            self._launch()  # warm-up: plans, workspace, kernel attributes
        torch.cuda.synchronize(dev)
        with torch.cuda.graph(self.graph, stream=self.stream, capture_error_mode="thread_local"):
            self._launch()
        torch.cuda.synchronize(dev)

    def _launch(self):
        self.ctx.call("psx_bk_raw", L.ptr(self.half), self.N, self.step, self.Ncut, self.Nmax,
                      L.ptr(self.braw), L.ptr(self.p0k), self.prec)

    def replay(self):
        """Run the captured sequence on the current stream (reads ``half``)."""
        self.graph.replay()

    def workspace_bytes(self):
        return self.ctx.workspace_bytes()

    def close(self):
        self.graph.reset()
        self.ctx.close()
