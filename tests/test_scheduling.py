"""Host-side scheduling logic of the projector pairs, on the CPU with stand-in
jobs: finish_each yields the pairs in index order whatever order their
truncation counts arrive in (a renormalisation's contraction order follows
from which of its two pairs comes first, so the order must not depend on
timing), and in eager mode the ready pairs right after the awaited one have
their products enqueued before anything is yielded."""
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
