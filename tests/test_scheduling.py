"""Host-side scheduling logic of the projector pairs, on the CPU with stand-in
jobs: finish_each yields the pairs in index order whatever order their
truncation counts arrive in (a renormalisation's contraction order follows
from which of its two pairs comes first, so the order must not depend on
timing), and in eager mode the ready pairs right after the awaited one have
their products enqueued before anything is yielded; start_many asks for the
range finder's fast path only after an ungraded spectrum of the same pair."""
import types

import pytest
import torch


class _Event:
    def __init__(self, *a, **k):
        pass

    def record(self, stream=None):
        pass


class _Ready:
    def __init__(self, log, idx, arrived):
        self.log, self.idx, self.arrived = log, idx, arrived

    def synchronize(self):
        self.log.append(("wait", self.idx))
        self.arrived[self.idx] = True

    def query(self):
        return self.arrived[self.idx]


class _Job:
    def __init__(self, log, idx, arrived):
        self.ready = _Ready(log, idx, arrived)
        self.graph = None
        self.spec = torch.zeros(4)
        self.spec_host = None
        self.key = idx
        self.stream = None
        self.log, self.idx = log, idx

    def finish(self):
        self.log.append(("finish", self.idx))
        return torch.zeros(8, 3), torch.zeros(8, 3)


@pytest.mark.parametrize("eager", [False, True])
@pytest.mark.parametrize("arrived", [[False, False, True, True], [False, True, False, True],
                                     [True, True, True, True], [False, False, False, True]])
def test_finish_each_yields_in_index_order(monkeypatch, eager, arrived):
    from acetn.renormalization.projectors import ProjectorCalculator
    monkeypatch.setattr(torch.cuda, "Event", _Event)
    calc = ProjectorCalculator.__new__(ProjectorCalculator)
    import collections
    calc.recent_spectra = collections.deque(maxlen=16)
    log = []
    initial = list(arrived)
    arrived = list(arrived)
    jobs = [_Job(log, i, arrived) for i in range(4)]
    net = types.SimpleNamespace(dims={'bond': 2})
    order = []
    for idx, (P1, P2), done in calc.finish_each(net, jobs, eager=eager):
        order.append(idx)
        log.append(("yield", idx))
        assert P1.shape == (2, 2, 2, 3)
    assert order == [0, 1, 2, 3]
    finished = [x for kind, x in log if kind == "finish"]
    assert finished == [0, 1, 2, 3]
    if eager:
        # the pairs right after the awaited one whose counts were already in
        # are finished before the first yield
        first_yield = log.index(("yield", 0))
        run = 1
        while run < 4 and initial[run]:
            run += 1
        assert [x for kind, x in log[:first_yield] if kind == "finish"] == list(range(run))
    else:
        for i in range(4):
            assert log.index(("finish", i)) < log.index(("yield", i))
            if i < 3:
                assert log.index(("yield", i)) < log.index(("finish", i + 1))


def test_range_finder_fast_path_follows_the_previous_spectrum(monkeypatch):
    """start_many asks for the range finder's fast path (RF_FAST) only for a
    pair whose previous spectrum was not graded (s_{q-1}/s_0 >= 1e-3): never
    on its first computation, and no longer once a spectrum came out graded."""
    from acetn import _hip
    from acetn.renormalization import projectors as pj
    monkeypatch.setattr(torch.cuda, "Event", _Event)
    made = []

    class FakeJob:
        def __init__(self, slot, sites, k, chi, p, niter, cutoff, seed, stream_id, **kw):
            self.rf_fast = kw["rf_fast"]
            self.graph = None
            self.ready = None
            self.stream = None
            self.spec = torch.zeros(4)
            self.spec_host = None
            made.append(self)

        def finish(self):
            return torch.zeros(8, 3), torch.zeros(8, 3)

    monkeypatch.setattr(_hip, "ProjectorJob", FakeJob)

    class Cfg:
        projectors = "half-system"
        svd_type = "rsvd"
        svd_cutoff = 1e-12
        rsvd_niter = 2
        rsvd_oversampling = 2
    calc = pj.ProjectorCalculator(Cfg())
    site = {'A': None, 'C': None, 'E': None}
    net = type("Net", (), {"dims": {'bond': 2, 'chi': 4},
                           "__getitem__": lambda self, s: site})()
    sites = [(0, 0), (1, 0), (1, 1), (0, 1)]

    def run(tail_ratio):
        """One computation of the pair whose spectrum ends at ``tail_ratio``
        s_0; returns whether it was asked for the fast path."""
        job = calc.start_many(net, [(sites, 0)])[0]
        job.spec_host = torch.tensor([1.0, 0.5, tail_ratio], dtype=torch.float64)
        list(calc.finish_each(net, [job]))
        return job.rf_fast

    assert run(1e-6) is False     # first computation: recovering path
    assert run(0.5) is False      # the previous spectrum was graded
    assert run(2e-3) is True      # the previous one was not
    assert run(5e-4) is True      # 2e-3 counts as not graded (>= 1e-3)
    assert run(0.5) is False      # 5e-4 is graded
    assert run(0.5) is True
