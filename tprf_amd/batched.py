"""Forward chi0 + RPA of many lattices with uneven (nk, norb, n_iw, beta) in one
HIP graph.

Every lattice keeps its own resident buffers (tprf_amd.chain.BubbleRPA).  One
eager pass fills the context's plans and tables and lets the library's kernel
autotuner pick a workgroup shape per shape class; the chains are then captured
into one graph, spread over K streams forked from and joined to the capture
stream, largest lattices first on the least-loaded stream.  ``autotune`` picks
K by timing each candidate graph.  A replay is one hipGraphLaunch for the whole
batch.
"""
import torch

from . import _lib
from .chain import BubbleRPA


def _cost(wl):
    """Algorithmic bytes of one forward eval (materialised tensors written once and
    read once), the weight used to balance the streams."""
    n, nk = wl.norb, wl.nk
    g = 2 * wl.n_iw * nk * n * n
    gt = wl.ntau * nk * n * n
    chi = wl.n_nu * nk * n ** 4
    return 16 * (5 * g + gt + 5 * chi)


class BatchedChain:
    def __init__(self, workloads, device="cuda"):
        self.dev = torch.device(device)
        self.chains = [BubbleRPA(wl, device=device) for wl in workloads]
        self.order = sorted(range(len(self.chains)), key=lambda i: -_cost(self.chains[i].wl))
        self.graph = None
        self.nstreams = None
        self._streams = []

    def run_eager(self):
        """Every lattice one after the other on the current stream (also the warm-up)."""
        for ch in self.chains:
            ch.run()

    def _assign(self, k):
        load, lanes = [0] * k, [[] for _ in range(k)]
        for i in self.order:
            j = min(range(k), key=load.__getitem__)
            lanes[j].append(i)
            load[j] += _cost(self.chains[i].wl)
        return lanes

    def capture(self, nstreams):
        """Capture the whole batch into one graph over ``nstreams`` streams."""
        self.run_eager()  # tables, plans and kernel variants are settled before capture
        torch.cuda.synchronize(self.dev)
        lanes = self._assign(nstreams)
        main = torch.cuda.Stream(device=self.dev)
        side = [torch.cuda.Stream(device=self.dev) for _ in range(nstreams - 1)]
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=main):
            fork = torch.cuda.Event()
            fork.record(main)
            done = []
            for j, lane in enumerate(lanes):
                st = main if j == 0 else side[j - 1]
                if j > 0:
                    st.wait_event(fork)
                with torch.cuda.stream(st):  # every stage binds the library context to this lane's stream
                    for i in lane:
                        for s in self.chains[i].STAGES:
                            self.chains[i].stage(s)
                if j > 0:
                    ev = torch.cuda.Event()
                    ev.record(st)
                    done.append(ev)
            for ev in done:
                main.wait_event(ev)
        _lib.ctx(self.dev.index)  # back to torch's current stream
        self.graph, self.nstreams, self._streams = g, nstreams, [main] + side
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
