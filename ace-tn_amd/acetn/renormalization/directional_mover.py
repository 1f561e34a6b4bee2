"""Directional moves (acetn/renormalization/directional_mover.py:5-366).

The move bookkeeping (which plaquette feeds which projector, which site
receives which renormalised tensor) is restated here on the host; the
arithmetic runs in acetn_hip_half_projectors / acetn_hip_renormalize_boundary.
"""
from .. import _hip
from .projectors import ProjectorCalculator


class DirectionalMover:
    def __init__(self, config, sketch=None):
        self.projector_calculator = ProjectorCalculator(config, sketch)
        self.calculate_projectors = self.projector_calculator.calculate

    # ---- moves: directional_mover.py:23-97
    def left_move(self, ipeps, xi):
        self.stage(ipeps, [(xi, 0)])

    def up_move(self, ipeps, yi):
        self.stage(ipeps, [(yi, 1)])

    def right_move(self, ipeps, xi):
        self.stage(ipeps, [(xi, 2)])

    def down_move(self, ipeps, yi):
        self.stage(ipeps, [(yi, 3)])

    def stage(self, ipeps, moves):
        """Run one or more moves whose projector computations and boundary
        writes are mutually independent -- a left/right (or up/down) pair of
        ctmrg.py:26-31 or fast_full_update.py:85-101: the left move of column
        x reads C0,C3,E0,E2,E3 of column x and writes C0,C3,E3 of column x+1,
        its right partner reads C1,C2,E0,E1,E2 and writes C1,C2,E1 -- so the
        result equals the sequential order.  Sketch streams are numbered in
        that sequential order.  With full-system projectors the right move
        reads corners the left move rewrites, so the moves run one after
        the other (each still with its projector pairs concurrent)."""
        if self.projector_calculator.projectors != "half-system" and len(moves) > 1:
            for m in moves:
                self._stage(ipeps, [m])
            return
        self._stage(ipeps, moves)

    def _stage(self, ipeps, moves):
        run_stages([(self, ipeps, moves)])

    def _plan(self, ipeps, moves):
        """Projector pairs (plaquette, k) of the moves and the renormalisations
        (s1, s2, i, j, k) that consume pairs i and j."""
        jobs, targets = [], []
        for line, k in moves:
            n_other = ipeps.ny if k in (0, 2) else ipeps.nx
            base = len(jobs)
            for idx in range(n_other):
                xi, yi = (line, idx) if k in (0, 2) else (idx, line)
                jobs.append((self.plaquette(ipeps.nx, ipeps.ny, xi, yi, k), k))
            for idx in range(n_other):
                s1, s2, i, j = self.targets(ipeps.nx, ipeps.ny, line, k, idx)
                targets.append((s1, s2, base + i, base + j, k))
        return jobs, targets

    def _move(self, ipeps, line, k):
        self.stage(ipeps, [(line, k)])

    # ---- calculate_{left,up,right,down}_projectors: directional_mover.py:99-181
    @staticmethod
    def plaquette(nx, ny, xi, yi, k):
        if k == 0:
            xj, yj = (xi + 1) % nx, (yi - 1 + ny) % ny
            return [(xi, yi), (xj, yi), (xj, yj), (xi, yj)]
        if k == 2:
            xj, yj = (xi - 1 + nx) % nx, (yi + 1) % ny
            return [(xi, yi), (xj, yi), (xj, yj), (xi, yj)]
        if k == 1:
            xj, yj = (xi - 1 + nx) % nx, (yi - 1 + ny) % ny
            return [(xi, yi), (xi, yj), (xj, yj), (xj, yi)]
        xj, yj = (xi + 1) % nx, (yi + 1) % ny
        return [(xi, yi), (xi, yj), (xj, yj), (xj, yi)]

    def calculate_left_projectors(self, ipeps, xi, yi):
        return self.calculate_projectors(ipeps, self.plaquette(ipeps.nx, ipeps.ny, xi, yi, 0), 0)

    def calculate_right_projectors(self, ipeps, xi, yi):
        return self.calculate_projectors(ipeps, self.plaquette(ipeps.nx, ipeps.ny, xi, yi, 2), 2)

    def calculate_up_projectors(self, ipeps, xi, yi):
        return self.calculate_projectors(ipeps, self.plaquette(ipeps.nx, ipeps.ny, xi, yi, 1), 1)

    def calculate_down_projectors(self, ipeps, xi, yi):
        return self.calculate_projectors(ipeps, self.plaquette(ipeps.nx, ipeps.ny, xi, yi, 3), 3)

    @staticmethod
    def targets(nx, ny, line, k, idx):
        """(s1, s2, i, j) of each renormalisation in a move."""
        if k == 0:
            return (line, idx), ((line + 1) % nx, idx), idx, (idx + 1) % ny
        if k == 2:
            return (line, idx), ((line - 1 + nx) % nx, idx), idx, (idx - 1 + ny) % ny
        if k == 1:
            return (idx, line), (idx, (line - 1 + ny) % ny), idx, (idx + 1) % nx
        return (idx, line), (idx, (line + 1) % ny), idx, (idx - 1 + nx) % nx

    def renormalize_boundary(self, ipeps, proj1, proj2, s1, s2, i, j, k):
        """directional_mover.py:273-303: writes C[(3+k)%4], C[k], E[(3+k)%4]
        of s2."""
        Ca, Cb, Eo = self._renormalize(ipeps, (proj1[i], proj2[i]), (proj1[j], proj2[j]), s1, k,
                                       precision=self.projector_calculator.precision)
        dst = ipeps[s2]
        dst['C'][(3 + k) % 4] = Ca
        dst['C'][k] = Cb
        dst['E'][(3 + k) % 4] = Eo

    @staticmethod
    def _renormalize(ipeps, pi, pj, s1, k, slot=None, precision="fp64", ws_key=None):
        st = ipeps[s1]
        flat = lambda P: P.reshape(-1, P.shape[-1])
        return _hip.renormalize_boundary((st['A'], st['C'], st['E']), k, flat(pi[0]), flat(pi[1]),
                                         flat(pj[0]), flat(pj[1]), slot=slot, precision=precision,
                                         ws_key=ws_key)


def run_stages(items, capture=None):
    """One stage of each of several independent networks, run as one round.

    ``items`` is a list of (mover, ipeps, moves).  Every projector pair of
    every item is started first (pair g of the round, counted over all
    items, on slot g % N_SLOTS with a workspace of its own, so that every
    stream used is one of the slot streams that waited on the caller's
    stream); the host then reads each pair's truncation
    count as soon as that pair has produced it, enqueues its projector
    products, and starts each renormalisation in two parts: the part that
    the first finished of its two pairs feeds on its own at once, the rest
    once the other pair is finished too (its slot stream waits on their
    completion events).  One
    network's Jacobi eigensolves thereby overlap the other pairs' matrix-core
    work.  The new boundary tensors are written back after everything has
    finished.  With full-system projectors the two moves of a stage depend on
    each other, so each network then runs its stage on its own.

    Under ``capture`` (a HIP graph capture of the caller's stream in
    progress, see graph.py) the same work is enqueued without any host
    synchronisation: the pairs take their truncation counts on the device
    and produce chi projector columns (those past the count zero), every
    buffer is kept alive by the capture, the workspaces are private to it,
    and the slot streams join the caller's stream at the end instead of a
    device synchronisation."""
    import torch
    if any(m.projector_calculator.projectors != "half-system" and len(mv) > 1
           for m, _, mv in items):
        if capture is not None:
            raise ValueError("graph capture: half-system projectors only")
        for m, net, mv in items:
            m.stage(net, mv)
        return
    plans, base, joined = [], 0, {}
    for mover, ipeps, moves in items:
        jobs, targets = mover._plan(ipeps, moves)
        calc = mover.projector_calculator
        if jobs:
            # the side streams first catch up with the caller's stream, where
            # the boundary tensors the pairs read may still be in flight
            dev = ipeps[tuple(jobs[0][0][0])]['A'].device
            if str(dev) not in joined:
                joined[str(dev)] = dev
                ev = torch.cuda.Event()
                ev.record(torch.cuda.current_stream(dev))
                for sl in range(1, _hip.N_SLOTS):
                    _hip.slot_stream(sl, dev).wait_event(ev)
        started = calc.start_many(ipeps, jobs, base=base, capture=capture)
        plans.append({"mover": mover, "ipeps": ipeps, "targets": targets, "started": started,
                      "base": base, "proj": [None] * len(jobs), "done": [None] * len(jobs),
                      "results": [None] * len(targets)})
        base += len(jobs)
    for pl in plans:
        mover, ipeps = pl["mover"], pl["ipeps"]
        calc = mover.projector_calculator
        proj, done, results = pl["proj"], pl["done"], pl["results"]
        parts = [None] * len(pl["targets"])
        flat = lambda P: P.reshape(-1, P.shape[-1])
        njobs = len(pl["started"])
        job_slot = [(pl["base"] + x) % _hip.N_SLOTS for x in range(njobs)]

        def clear(sl):
            # every pair of this network on slot sl has its projector
            # products enqueued
            return all(proj[x] is not None for x in range(njobs) if job_slot[x] == sl)

        for idx, pr, ev in calc.finish_each(ipeps, pl["started"], eager=True):
            proj[idx], done[idx] = pr, ev
            for t, (s1, s2, i, j, k) in enumerate(pl["targets"]):
                if results[t] is not None:
                    continue
                g = pl["base"] + t
                home = g % _hip.N_SLOTS
                dev = ipeps[s1]['A'].device
                # the part that the first finished of the two pairs feeds on
                # its own starts at once (on a stream whose pairs are past
                # their eigensolvers, so that it holds up no projector
                # products queued there later), the rest when the other pair
                # is done too (same workspace)
                if parts[t] is None and (proj[j] is not None or proj[i] is not None):
                    a = j if proj[j] is not None else i
                    sl = home if clear(home) else job_slot[a] if clear(job_slot[a]) else home
                    stream = _hip.slot_stream(sl, dev)
                    stream.wait_event(done[a])
                    with torch.cuda.stream(stream):
                        st = ipeps[s1]
                        parts[t] = [_hip.RenormalizeParts(
                            (st['A'], st['C'], st['E']), k, ipeps.dims['chi'],
                            first_pair="j" if a == j else "i",
                            slot=None if capture is not None else home,
                            precision=calc.precision,
                            ws_key=None if capture is not None else g), a, sl, None]
                        parts[t][0].first(flat(proj[a][0]), flat(proj[a][1]))
                        parts[t][3] = torch.cuda.Event()
                        parts[t][3].record(stream)
                if parts[t] is not None and proj[i] is not None and proj[j] is not None:
                    b = i if parts[t][1] == j else j
                    sl = home if clear(home) else parts[t][2]
                    stream = _hip.slot_stream(sl, dev)
                    stream.wait_event(done[b])
                    if sl != parts[t][2]:
                        stream.wait_event(parts[t][3])   # the first part (same workspace)
                    with torch.cuda.stream(stream):
                        results[t] = parts[t][0].second(flat(proj[b][0]), flat(proj[b][1]))
        pl["parts"] = parts
    if capture is None:
        torch.cuda.synchronize()
    else:
        for dev in joined.values():
            cur = torch.cuda.current_stream(dev)
            for sl in range(1, _hip.N_SLOTS):
                cur.wait_stream(_hip.slot_stream(sl, dev))
        capture.keep.append(plans)
    for pl in plans:
        for (s1, s2, i, j, k), (Ca, Cb, Eo) in zip(pl["targets"], pl["results"]):
            dst = pl["ipeps"][s2]
            dst['C'][(3 + k) % 4] = Ca
            dst['C'][k] = Cb
            dst['E'][(3 + k) % 4] = Eo


def run_chain_pair(mover, ipeps, chain_a, chain_b, stagger=True):
    """The two independent move chains of one half of a CTMRG iteration,
    pipelined.  ``chain_a`` / ``chain_b`` are the left (up) and right (down)
    moves of ctmrg_iteration's stages, in stage order: the left move of
    column x reads C0, C3, E0, E2, E3 of column x and writes C0, C3, E3 of
    column x+1, the right move reads C1, C2, E0, E1, E2 and writes C1, C2, E1
    (likewise up/down), so each chain depends only on its own previous move
    and the two chains share nothing they write.

    Stage k of chain a runs on slots 0-1, of chain b on slots 2-3.  Chain b's
    pairs of stage k start their matrix-core work once chain a's pairs of that
    stage have enqueued theirs (their mark events, recorded ahead of the
    one-workgroup eigensolver), and chain a's pairs of stage k+1 once chain b's
    of stage k have: each chain's latency-bound eigensolves run while the
    other chain's products fill the GPU, where running the two moves of a
    stage side by side puts all four eigensolves at the same time.  Every pair
    and renormalisation computes exactly what run_stages computes (same
    kernels, inputs and sketch streams, taken in the same order), so the
    results are the same bit for bit.  Workspaces are keyed per chain, stage
    and pair (each used once per call), and the tensors a chain replaces stay
    referenced until the final synchronisation."""
    import torch
    calc = mover.projector_calculator
    n = len(chain_a)
    if len(chain_b) != n:
        raise ValueError("run_chain_pair: chains of different length")
    chains = (chain_a, chain_b)
    slots = ((0, 1), (2, 3))
    # the device of the first plaquette's sites (a network needs no site
    # list here, only indexing by site)
    dev = ipeps[tuple(mover._plan(ipeps, [chain_a[0]])[0][0][0][0])]['A'].device
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(dev))
    for sl in range(1, _hip.N_SLOTS):
        _hip.slot_stream(sl, dev).wait_event(ev)
    keep = []
    renorm_done = [[], []]   # completion events of each chain's latest renormalisations
    marks = [[], []]         # mark events of each chain's latest started pairs
    state = [None, None]     # (jobs, targets, started) of each chain's stage in flight

    def begin(c, k):
        jobs, targets = mover._plan(ipeps, [chains[c][k]])
        for sl in slots[c]:
            st = _hip.slot_stream(sl, dev)
            for e in renorm_done[c] + (marks[1 - c] if stagger else []):
                st.wait_event(e)
        js = calc.start_many(ipeps, jobs, slots=[slots[c][i % 2] for i in range(len(jobs))],
                             ws_keys=[("chain", c, k, i) for i in range(len(jobs))],
                             mark=stagger)
        marks[c] = [j.mark for j in js] if stagger else []
        state[c] = (targets, js, k)
        keep.append(js)

    def finish(c):
        targets, js, k = state[c]
        proj, done = [None] * len(js), [None] * len(js)
        parts, results = [None] * len(targets), [None] * len(targets)
        flat = lambda P: P.reshape(-1, P.shape[-1])
        ends = []
        job_slot = [slots[c][x % 2] for x in range(len(js))]

        def clear(sl):
            # every pair on slot sl has its projector products enqueued, so
            # work added there now runs behind nothing but them
            return all(proj[x] is not None for x in range(len(js)) if job_slot[x] == sl)

        for idx, pr, dn in calc.finish_each(ipeps, js, eager=True):
            proj[idx], done[idx] = pr, dn
            for t, (s1, s2, i, j, kk) in enumerate(targets):
                if results[t] is not None:
                    continue
                home = slots[c][t % 2]
                if parts[t] is None and (proj[j] is not None or proj[i] is not None):
                    a = j if proj[j] is not None else i
                    # a part placed on a stream whose pair is still in its
                    # eigensolver would hold up that pair's projector products,
                    # queued behind it later: it goes to pair a's stream then
                    sl = home if clear(home) else job_slot[a] if clear(job_slot[a]) else home
                    stream = _hip.slot_stream(sl, dev)
                    stream.wait_event(done[a])
                    with torch.cuda.stream(stream):
                        st = ipeps[s1]
                        parts[t] = [_hip.RenormalizeParts(
                            (st['A'], st['C'], st['E']), kk, ipeps.dims['chi'],
                            first_pair="j" if a == j else "i", slot=sl,
                            precision=calc.precision, ws_key=("chain", c, k, t)), a, sl, None]
                        parts[t][0].first(flat(proj[a][0]), flat(proj[a][1]))
                        parts[t][3] = torch.cuda.Event()
                        parts[t][3].record(stream)
                if parts[t] is not None and proj[i] is not None and proj[j] is not None:
                    b = i if parts[t][1] == j else j
                    sl = home if clear(home) else parts[t][2]
                    stream = _hip.slot_stream(sl, dev)
                    stream.wait_event(done[b])
                    if sl != parts[t][2]:
                        stream.wait_event(parts[t][3])   # the first part (same workspace)
                    with torch.cuda.stream(stream):
                        results[t] = parts[t][0].second(flat(proj[b][0]), flat(proj[b][1]))
                        e = torch.cuda.Event()
                        e.record(stream)
                        ends.append(e)
        keep.append((proj, parts, results))
        renorm_done[c] = ends
        # the next move of this chain reads what this one wrote
        for (s1, s2, i, j, kk), (Ca, Cb, Eo) in zip(targets, results):
            dst = ipeps[s2]
            keep.append((dst['C'][(3 + kk) % 4], dst['C'][kk], dst['E'][(3 + kk) % 4]))
            dst['C'][(3 + kk) % 4] = Ca
            dst['C'][kk] = Cb
            dst['E'][(3 + kk) % 4] = Eo

    begin(0, 0)
    begin(1, 0)
    for k in range(n):
        finish(0)
        if k + 1 < n:
            begin(0, k + 1)
        finish(1)
        if k + 1 < n:
            begin(1, k + 1)
    torch.cuda.synchronize()
    del keep
