"""One CTMRG iteration captured in a HIP graph and replayed.

The reference's loop (acetn/renormalization/ctmrg.py:26-31) reads every
projector pair's truncation count cD_new = min(chi, #{s/s0 > cutoff}) on the
host (projectors.py:163-164), so the shapes of the next products are only
known after a synchronisation.  Once every boundary dimension of a network
has saturated at chi, this module instead enqueues the iteration with fixed
shapes: each pair keeps its count on the device and writes chi projector
columns, the columns at and past the count zero -- an exact embedding of the
truncated projectors, so a renormalised tensor carries zero rows and columns
where the reference's would be shorter.  With no host synchronisation left
inside, the whole iteration (all stages, all four slot streams, one or
several networks as in ctmrg_iteration_many) is captured once in a
torch.cuda.CUDAGraph and replayed.

State: the graph reads and writes fixed buffers.  Capture replaces each
network's C and E tensors by contiguous copies that the graph owns, and the
iteration's last step copies the new tensors back into them, so successive
replays chain on the device.  Sketches: pair i of an iteration draws Philox
stream (counter + i), the counter a device value set from the network's
SketchCounter before every replay and advanced on the host after it -- the
same streams an uncaptured iteration draws.  With the same kernel choices
(tile tables included) a replay therefore reproduces an uncaptured iteration
bit for bit.  The counts of a replay are read with ``counts_ok``; a count
below chi (a boundary that lost rank) leaves a valid padded state, and the
caller continues uncaptured, where the next iteration truncates the zero
directions.

Status: switched off.  The first capture of a multi-stream iteration on the
MI355X (D=3, chi=12, FP64, four slot streams forked from and joined back to
the capture stream at every stage) ended in a segmentation fault inside
torch.cuda.CUDAGraph.capture_end, i.e. in the runtime's end-of-capture or
graph instantiation, after every library call had returned success.  No
mistake in this code has been found that explains it, so CapturedIteration
raises until one is; the fixed-shape projector entry points it uses
(acetn_hip_projectors_begin_graph / _end_padded) run uncaptured and are
tested that way.
"""
import torch

from .ctmrg import stage_moves

CAPTURE_ENABLED = False
_DISABLED = ("graph capture of the CTMRG iteration is switched off: its first run ended in a "
             "segmentation fault at the end of stream capture, with no cause found yet")


def saturated(ipeps):
    """Whether every corner is chi x chi and every edge chi x chi x D x D."""
    chi = ipeps.dims['chi']
    for s in ipeps.site_list:
        st = ipeps[s]
        if any(tuple(c.shape) != (chi, chi) for c in st['C']):
            return False
        if any(tuple(e.shape[:2]) != (chi, chi) for e in st['E']):
            return False
    return True


class _Capture:
    """Bookkeeping while the iteration is being captured: sketch-stream
    offsets per network, the device count slots, every buffer to keep."""

    def __init__(self, calcs, base, counts):
        self.index = {id(c): i for i, c in enumerate(calcs)}
        self.offset = [0] * len(calcs)
        self.base, self.counts = base, counts
        self.pair = 0
        self.keep = []

    def next_pair(self, calc):
        i = self.index[id(calc)]
        sid = self.offset[i]
        self.offset[i] += 1
        g = self.pair
        self.pair += 1
        if g >= self.counts.numel():
            raise RuntimeError("graph capture: more projector pairs than planned")
        return sid, (self.base[i:i + 1], self.counts, g)


class CapturedIteration:
    """``pairs``: list of (mover, ipeps), each network saturated, half-system
    rSVD projectors.  Construction captures one iteration of all of them
    (the round order of ctmrg_iteration_many; for one network that is
    ctmrg_iteration's order) without running it; ``replay()`` runs it."""

    def __init__(self, pairs):
        from .directional_mover import run_stages
        self.pairs = list(pairs)
        if not self.pairs:
            raise ValueError("CapturedIteration: no network")
        for mover, net in self.pairs:
            calc = mover.projector_calculator
            if calc.projectors != "half-system" or calc.svd_type != "rsvd":
                raise ValueError("graph capture: half-system rsvd projectors only")
            if not saturated(net):
                raise ValueError("graph capture: the boundary dimension has not saturated at chi")
        self.calcs = [m.projector_calculator for m, _ in self.pairs]
        if len(set(map(id, self.calcs))) != len(self.calcs):
            raise ValueError("graph capture: each network needs a mover of its own")
        if not CAPTURE_ENABLED:
            raise NotImplementedError(_DISABLED)
        dev = self.pairs[0][1][self.pairs[0][1].site_list[0]]['A'].device
        self.device = dev
        plans = [(m, net, stage_moves(net)) for m, net in self.pairs]
        self.npairs = [0] * len(self.pairs)
        for i, (m, net, stages) in enumerate(plans):
            for moves in stages:
                jobs, _ = m._plan(net, moves)
                self.npairs[i] += len(jobs)
        self.stream_base = torch.zeros(len(self.pairs), dtype=torch.int64, device=dev)
        self.counts = torch.zeros(sum(self.npairs), dtype=torch.int32, device=dev)
        # fixed state buffers, owned here
        self.state = []
        for _, net in self.pairs:
            st = {}
            for s in net.site_list:
                C, E = net[s]['C'], net[s]['E']  # the site's own lists
                for k in range(4):
                    C[k] = C[k].contiguous().clone()
                    E[k] = E[k].contiguous().clone()
                st[s] = (list(C), list(E))
            self.state.append(st)
        torch.cuda.synchronize(dev)
        cap = _Capture(self.calcs, self.stream_base, self.counts)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for s in range(max(len(st) for _, _, st in plans)):
                run_stages([(m, net, st[s]) for m, net, st in plans if s < len(st)],
                           capture=cap)
            for (_, net), st in zip(self.pairs, self.state):
                for s, (C, E) in st.items():
                    site = net[s]
                    for k in range(4):
                        if site['C'][k] is not C[k]:
                            C[k].copy_(site['C'][k])
                        if site['E'][k] is not E[k]:
                            E[k].copy_(site['E'][k])
        if cap.offset != self.npairs:
            raise RuntimeError("graph capture: projector pair count differs from the plan")
        self.keep = cap.keep
        # the networks read and write the fixed buffers from now on
        for (_, net), st in zip(self.pairs, self.state):
            for s, (C, E) in st.items():
                for k in range(4):
                    net[s]['C'][k] = C[k]
                    net[s]['E'][k] = E[k]
        self.replays = 0

    def replay(self):
        """One iteration of every network, enqueued on the current stream."""
        for i, calc in enumerate(self.calcs):
            self.stream_base[i].fill_(calc.sketch.next)
        self.graph.replay()
        for calc, n in zip(self.calcs, self.npairs):
            calc.sketch.next += n
        self.replays += 1

    def counts_ok(self):
        """Synchronises; whether every pair of the last replay kept chi."""
        chis = []
        for (_, net), n in zip(self.pairs, self.npairs):
            chis += [net.dims['chi']] * n
        got = self.counts.cpu()
        return bool((got == torch.tensor(chis, dtype=got.dtype)).all())
