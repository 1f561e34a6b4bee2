"""The CTMRG iteration captured in a HIP graph (acetn/renormalization/graph.py)
against the uncaptured iteration (ctmrg.py:26-31 restated by ctmrg_iteration):
bit for bit once the boundary dimension has saturated, for one network and
for several uneven ones together (ctmrg_iteration_many); the fixed-shape
projector pair (device truncation count, chi columns with those past the
count zero) against the host-count pair; the config switch end to end."""
import numpy as np
import pytest
import torch

from oracle import ctmrg as oc
from test_hip_ctmrg import Cfg, DevState, to_dev_site

pytestmark = pytest.mark.gpu

# The capture itself is switched off (acetn.renormalization.graph): the first
# capture of the four-stream iteration ended in a segmentation fault inside
# CUDAGraph.capture_end with no cause found in this code.
CAPTURE_OFF = pytest.mark.skip(reason="graph capture switched off: segmentation fault at the "
                                      "end of stream capture, cause not found")


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def assert_states_equal(a, b, tag):
    for s in a.site_list:
        for key in ('C', 'E'):
            for k in range(4):
                x, y = a[s][key][k], b[s][key][k]
                assert x.shape == y.shape, (tag, s, key, k, x.shape, y.shape)
                assert torch.equal(x, y), (tag, s, key, k, float((x - y).abs().max()))


def _pair(D, chi, seed, precision):
    from acetn.renormalization.directional_mover import DirectionalMover
    st = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=500 + seed))
    cfg = Cfg()
    cfg.precision = precision
    nets = []
    for _ in range(2):
        net = DevState(st, 2, 2, D, chi, seed)
        nets.append((DirectionalMover(cfg, net.sketch), net))
    return nets


@CAPTURE_OFF
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("D,chi", [(3, 12), (4, 32)])
def test_captured_iteration_equals_uncaptured(hip, precision, D, chi):
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.graph import CapturedIteration, saturated
    (ma, a), (mb, b) = _pair(D, chi, D + chi, precision)
    for _ in range(2):  # saturate chi, same work on both
        ctmrg_iteration(ma, a)
        ctmrg_iteration(mb, b)
    assert saturated(a) and saturated(b)
    assert_states_equal(a, b, "warmup")
    g = CapturedIteration([(mb, b)])
    for it in range(3):
        ctmrg_iteration(ma, a)
        g.replay()
        torch.cuda.synchronize()
        assert g.counts_ok()
        assert a.sketch.next == b.sketch.next
        assert_states_equal(a, b, f"replay {it}")
    print(f"D={D} chi={chi} {precision}: 3 replays bit-equal to the uncaptured iterations")


@CAPTURE_OFF
def test_captured_many_networks_equal_uncaptured(hip):
    """Uneven instances (the parameter-sweep case) captured in one graph."""
    from acetn.renormalization.ctmrg import ctmrg_iteration_many
    from acetn.renormalization.graph import CapturedIteration
    inst = [(2, 8, 1), (3, 12, 2), (4, 16, 3)]
    pairs = [_pair(D, chi, s, "fp64") for D, chi, s in inst]
    eager = [p[0] for p in pairs]
    capt = [p[1] for p in pairs]
    for _ in range(2):
        ctmrg_iteration_many(eager)
        ctmrg_iteration_many(capt)
    g = CapturedIteration(capt)
    for it in range(2):
        ctmrg_iteration_many(eager)
        g.replay()
        torch.cuda.synchronize()
        assert g.counts_ok()
        for (_, a), (_, b), i in zip(eager, capt, inst):
            assert_states_equal(a, b, f"instance {i} replay {it}")


def test_capture_refuses_unsaturated_network(hip):
    from acetn.renormalization.graph import CapturedIteration
    (m, net), _ = _pair(3, 12, 7, "fp64")
    with pytest.raises(ValueError, match="saturated"):
        CapturedIteration([(m, net)])


@pytest.mark.parametrize("cutoff", [1e-12, 3e-2])
def test_padded_projectors_embed_the_truncated_ones(hip, cutoff):
    """acetn_hip_projectors_begin_graph / _end_padded against _begin / _end on
    the same plaquette and sketch: the count on the device equals the host
    count, the first cD columns are bit-equal, the rest are zero.  The large
    cutoff makes the count fall below chi."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    (m, net), _ = _pair(3, 16, 9, "fp64")
    for _ in range(2):
        ctmrg_iteration(m, net)
    sites = [tuple(s) for s in oc.plaquette(2, 2, 0, 1, 0)]
    site_t = [(net[s]['A'], net[s]['C'], net[s]['E']) for s in sites]
    chi, stream_id = 16, 41
    ref = hip.ProjectorJob(0, site_t, 0, chi, 2, 2, cutoff, 123, stream_id)
    torch.cuda.synchronize()
    P1r, P2r = ref.finish()
    base = torch.full((1,), stream_id - 5, dtype=torch.int64, device="cuda")
    counts = torch.zeros(3, dtype=torch.int32, device="cuda")
    job = hip.ProjectorJob(0, site_t, 0, chi, 2, 2, cutoff, 123, 5, graph=(base, counts, 1))
    P1p, P2p = job.finish()
    torch.cuda.synchronize()
    cd = P1r.shape[1]
    print(f"cutoff {cutoff}: kept {cd} of chi={chi}")
    if cutoff > 1e-6:
        assert cd < chi
    assert int(counts[1]) == cd and int(counts[0]) == 0 and int(counts[2]) == 0
    assert P1p.shape == (P1r.shape[0], chi)
    assert torch.equal(P1p[:, :cd], P1r) and torch.equal(P2p[:, :cd], P2r)
    assert not P1p[:, cd:].any() and not P2p[:, cd:].any()
    assert torch.equal(job.spec, ref.spec)


@CAPTURE_OFF
def test_ctmrg_graph_config_matches_uncaptured(hip):
    """ctmrg(config.graph=True) through the Ipeps API: the replays give the
    same tensors as the uncaptured run from the same state and seed."""
    from acetn.ipeps import Ipeps
    out = []
    for graph in (False, True):
        cfg = {'dtype': 'float64', 'device': 'cuda',
               'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': 3, 'chi': 12}},
               'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
               'ctmrg': {'steps': 6, 'disable_progressbar': True, 'sketch_seed': 17,
                         'graph': graph}}
        ipeps = Ipeps(cfg)
        A = oc.random_site_tensors(2, 2, 3, 2, seed=77)
        for site in ipeps.site_list:
            ipeps[site]['A'] = torch.as_tensor(A[tuple(site)])
            ipeps[site].initialize_corner_tensors()
            ipeps[site].initialize_edge_tensors()
        ipeps.site_states_initialized = True
        ipeps.renormalize()
        out.append((ipeps, float(ipeps.measure()['Energy'])))
    (a, ea), (b, eb) = out
    assert ea == eb
    for s in a.site_list:
        for key in ('C', 'E'):
            for k in range(4):
                assert torch.equal(a[s][key][k], b[s][key][k]), (s, key, k)

