"""Projector construction (acetn/renormalization/projectors.py:6-236) on the
HIP engine.  The enlarged corners, the randomized range finder, the FP64
Jacobi SVD and the projector products all run on the GPU
(acetn_hip_half_projectors)."""
import collections

from .. import _hip


# s_{q-1} / s_0 at or above which a pair's spectrum counts as not graded
UNGRADED = 1e-3


class SketchCounter:
    """Numbers the rSVD calls: call i uses Philox stream (first + i) under
    ``seed``; the reference draws a fresh torch.randn sketch per call
    (acetn/linalg/fused_matmul_svd_lowrank.py:32)."""

    def __init__(self, seed, first=0):
        self.seed = int(seed)
        self.next = int(first)

    def take(self):
        s = self.next
        self.next += 1
        return s


class ProjectorCalculator:
    def __init__(self, config, sketch=None):
        self.projectors = config.projectors
        self.svd_type = config.svd_type
        self.svd_cutoff = config.svd_cutoff
        self.rsvd_niter = config.rsvd_niter
        self.rsvd_oversampling = config.rsvd_oversampling
        self.sketch = sketch if sketch is not None else SketchCounter(0)
        self.precision = getattr(config, "precision", "fp64")
        _hip.precision_code(self.precision)  # validates
        self.last_spectrum = None
        # (spectrum, kept count) of the most recent projector pairs, newest
        # last: the singular values the truncation chooses from
        self.recent_spectra = collections.deque(maxlen=256)
        # pair (sites, k, system) -> its last spectrum was not graded, so the
        # next computation may take the range finder's fast path (RF_FAST)
        self.ungraded = {}
        self.set_calculate()

    def set_calculate(self):
        """projectors.py:21-33."""
        if self.projectors is None or self.projectors == "full-system":
            self.calculate = self.calculate_full_system
        elif self.projectors == "half-system":
            self.calculate = self.calculate_half_system
        else:
            raise ValueError(f"Invalid ctmrg projector type: {self.projectors} provided.")

    @staticmethod
    def _site(ipeps, s):
        st = ipeps[s]
        return st['A'], st['C'], st['E']

    def calculate_half_system(self, ipeps, sites, k):
        """projectors.py:138-174 (rsvd) / :62-82 + :114-136 (full-rank)."""
        return self.calculate_many(ipeps, [(sites, k)], system="half-system")[0]

    def calculate_full_system(self, ipeps, sites, k):
        """projectors.py:176-217 (rsvd) / :84-112 + :114-136 (full-rank)."""
        return self.calculate_many(ipeps, [(sites, k)], system="full-system")[0]

    def calculate_many(self, ipeps, jobs, system=None):
        """Projector pairs of mutually independent plaquettes, ``jobs`` a list
        of (sites, k) in the reference's call order (which fixes their sketch
        streams).  Each pair runs on its own stream; the host reads each
        truncation rank as soon as its pair has computed it."""
        import torch
        out = list(self.finish_each(ipeps, self.start_many(ipeps, jobs, system)))
        torch.cuda.synchronize()
        return [pr for _, pr, _ in out]

    def start_many(self, ipeps, jobs, system=None, base=0, capture=None, slots=None,
                   ws_keys=None, mark=False):
        """Enqueue every projector pair up to its truncation count; returns
        the started jobs in ``jobs`` order.  Pair i is round-wide pair
        base + i (``base`` counts the pairs of networks started before it in
        the same round) and runs on slot (base + i) % N_SLOTS with a
        workspace of its own.  Under ``capture`` (a graph capture in
        progress) the pairs are enqueued in the capturable form, their
        sketch streams numbered by the capture.  ``slots`` / ``ws_keys``
        (one per job) override the slot and workspace assignment; ``mark``
        records each pair's mark event (ProjectorJob)."""
        system = system or ("full-system" if self.calculate == self.calculate_full_system
                            else "half-system")
        if self.svd_type not in ("rsvd", "full-rank"):
            raise ValueError(f"Invalid svd_type: {self.svd_type}")
        chi = ipeps.dims['chi']
        started = []
        for i, (sites, k) in enumerate(jobs):
            slot = (base + i) % _hip.N_SLOTS if slots is None else slots[i]
            ws_key = base + i if ws_keys is None else ws_keys[i]
            sites = [tuple(s) for s in sites]
            for s in sites:
                ipeps[s]  # raises ValueError for an unknown site
            graph = None
            if capture is not None:
                stream_id, graph = capture.next_pair(self)
            else:
                stream_id = self.sketch.take() if self.svd_type == "rsvd" else 0
            key = (tuple(sites), k % 4, system)
            job = _hip.ProjectorJob(slot, [self._site(ipeps, s) for s in sites], k % 4,
                                    chi, self.rsvd_oversampling, self.rsvd_niter,
                                    self.svd_cutoff, self.sketch.seed, stream_id,
                                    precision=self.precision, system=system,
                                    svd=self.svd_type, ws_key=ws_key, graph=graph,
                                    mark=mark,
                                    rf_fast=graph is None and self.ungraded.get(key, False))
            job.key = key
            started.append(job)
        return started

    def finish_each(self, ipeps, started, eager=False):
        """Yields (index, (P1, P2), done_event) per started job, in index
        order: the host waits for that job's truncation count only, then
        enqueues its projector products on the job's stream; ``done_event``
        marks their completion on that stream.  With ``eager``, once a job's
        count is in, the jobs right after it whose counts have arrived too
        have their products enqueued before any of them is yielded, so that
        work the caller enqueues per yielded job is never queued ahead of
        those products on their streams.  The order stays the index order
        (callers pick the contraction order of a renormalisation from which
        of its two pairs comes first, so a timing-dependent order would make
        the results depend on timing at the rounding level)."""
        import torch
        D = ipeps.dims['bond']

        def finish(job):
            if job.spec_host is not None:
                # the range finder's CholeskyQR keeps every column when the
                # q values span less than 1e-3 (margin 100 on its rank guard)
                s = job.spec_host
                self.ungraded[job.key] = bool(s[0] > 0 and s[-1] >= UNGRADED * s[0])
            P1, P2 = job.finish()
            # the kept count: a device scalar for a captured pair (written on
            # every replay)
            cd = P1.shape[1] if job.graph is None else job.graph[1][job.graph[2]]
            self.recent_spectra.append((job.spec, cd))
            done = torch.cuda.Event()
            done.record(job.stream)
            rows = P1.shape[0]
            return (P1.view(rows // (D * D), D, D, -1), P2.view(rows // (D * D), D, D, -1)), done

        pending = list(range(len(started)))
        while pending:
            idx = pending.pop(0)
            if started[idx].ready is not None:
                started[idx].ready.synchronize()
            batch = [idx]
            while eager and pending and (started[pending[0]].ready is None or
                                         started[pending[0]].ready.query()):
                batch.append(pending.pop(0))
            outs = [(x, finish(started[x])) for x in batch]
            for x, (P, done) in outs:
                yield x, P, done
        if started:
            self.last_spectrum = started[-1].spec

    def truncation_errors(self, count):
        """Projector truncation error of the ``count`` most recent pairs: the
        share of the computed spectrum's weight that the truncation drops,
        sum_{i >= cD} s_i^2 / sum_i s_i^2 over the q = chi + oversampling
        singular values of the half-system matrix (the reference keeps
        cD = min(chi, #{s/s0 > cutoff}) of them, projectors.py:163-164, and
        computes no error)."""
        import torch
        out = []
        for spec, cd in list(self.recent_spectra)[-count:]:
            cd = int(cd)
            s2 = spec.double().cpu() ** 2
            tot = float(s2.sum())
            out.append(float(s2[cd:].sum()) / tot if tot > 0 else 0.0)
        return out

