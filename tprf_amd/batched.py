"""Forward chi0 + RPA of many lattices with uneven (nk, norb, n_iw, beta) in one
call of the C entry tprf_forward_batched, captured into one HIP graph.

Every lattice keeps its own resident buffers (tprf_amd.chain.BubbleRPA); a
table of per-instance descriptors and device pointers drives the library,
which runs the instances on several streams forked from and joined to the
caller's stream, largest first on the least-loaded stream.  One eager call
fills the context's plans and tables and lets the kernel autotuner settle a
workgroup shape per shape class; the call is then captured, and a replay is
one hipGraphLaunch for the whole batch.  ``autotune`` picks the stream count by
timing each candidate graph.
"""
import ctypes

import torch

from . import _lib
from .chain import BubbleRPA


class BatchedChain:
    def __init__(self, workloads, device="cuda"):
        self.dev = torch.device(device)
        if self.dev.index is None:
            self.dev = torch.device(self.dev.type, torch.cuda.current_device())
        self.chains = [BubbleRPA(wl, device=self.dev) for wl in workloads]
        n = len(self.chains)
        self._desc = (_lib.ChainDesc * n)()
        self._ptrs = (_lib.ChainPtrs * n)()
        for i, ch in enumerate(self.chains):
            wl, d, p = ch.wl, self._desc[i], self._ptrs[i]
            d.beta, d.n_iw, d.ntau, d.nb, d.norb, d.mu = wl.beta, wl.n_iw, wl.ntau, wl.nb, wl.norb, wl.mu
            for k in range(3):
                d.dims[k] = ch.dims[k]
            d.flags = (_lib.TPRF_CHAIN_HERMITIAN_EK if ch.herm else 0) | (_lib.TPRF_CHAIN_HERMITIAN_U if ch.herm_nu else 0)
            for f, t in (("e_k", ch.e_k), ("U", ch.U), ("g_wr", ch.g_wr), ("g_tr", ch.g_tr),
                         ("chi0_wr", ch._chi0_wr), ("chi0_wk", ch._chi0_wk), ("chi_wk", ch.chi_wk)):
                setattr(p, f, t.data_ptr())
            # the library's size rule must agree with the buffers the chain holds
            counts = (ctypes.c_long * 5)()
            _lib.check(_lib.lib().tprf_chain_buffer_sizes(ctypes.byref(d), counts), None, "chain buffer sizes")
            got = [ch.g_wr.numel(), ch.g_tr.numel(), ch._chi0_wr.numel(), ch._chi0_wk.numel(), ch.chi_wk.numel()]
            if list(counts) != got:
                raise RuntimeError(f"instance {i}: buffer sizes {got} != library's {list(counts)}")
        self.graph = None
        self.nstreams = None

    def run_batched(self, nstreams=1):
        """The whole batch as one library call on torch's current stream."""
        c = _lib.ctx(self.dev.index)
        _lib.call("tprf_forward_batched", c, len(self.chains), ctypes.byref(self._desc), ctypes.byref(self._ptrs),
                  int(nstreams))

    def run_eager(self):
        """Every lattice one after the other on the current stream, one library call
        per stage: the per-instance loop of the same chain."""
        for ch in self.chains:
            ch.run()

    def capture(self, nstreams):
        """Capture the whole batch into one graph over ``nstreams`` streams."""
        self.run_batched(1)  # tables, plans and kernel variants are settled on one stream
        self.run_batched(nstreams)  # the library's lane streams and events exist before capture
        torch.cuda.synchronize(self.dev)
        main = torch.cuda.Stream(device=self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=main):
            self.run_batched(nstreams)
        _lib.ctx(self.dev.index)  # back to torch's current stream
        self.graph, self.nstreams = g, nstreams
        return g

    def replay(self):
        self.graph.replay()

    def autotune(self, candidates=(1, 2, 4, 8), reps=3):
        """Capture one graph per stream count, time ``reps`` replays of each and keep
        the fastest.  Returns {nstreams: ms per batch}."""
        times = {}
        best = None
        for k in candidates:
            self.capture(k)
            self.replay()
            torch.cuda.synchronize(self.dev)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                self.replay()
            e1.record()
            torch.cuda.synchronize(self.dev)
            times[k] = e0.elapsed_time(e1) / reps
            if best is None or times[k] < times[best]:
                best = k
        if best != self.nstreams:
            self.capture(best)
        return times

    def check_status(self):
        """(max tail-fit residual, max |m0|) over the batch since the last reset."""
        return _lib.tail_status(_lib.ctx(self.dev.index))

    def results(self):
        return [ch.chi_wk for ch in self.chains]
