"""CTMRG driver (acetn/renormalization/ctmrg.py:4-31)."""
import torch
from tqdm import tqdm

from .directional_mover import DirectionalMover


def corner_spectra(ipeps):
    out = []
    for s in ipeps.site_list:
        for c in ipeps[s]['C']:
            sv = torch.linalg.svdvals(c.cpu())
            out.append(sv / sv[0])
    return out


def _spectra_change(a, b):
    m = 0.0
    for x, y in zip(a, b):
        n = max(len(x), len(y))
        xx = torch.zeros(n, dtype=x.dtype); yy = torch.zeros(n, dtype=y.dtype)
        xx[:len(x)] = x; yy[:len(y)] = y
        m = max(m, float((xx - yy).abs().max()))
    return m


def ctmrg_iteration(mover, ipeps, pipelined="auto"):
    """One pass of ctmrg.py:26-31 (non-distributed order).  Each
    left/right (up/down) pair is independent and runs as one stage; with
    half-system projectors the left and right (up and down) moves form two
    independent chains, which run pipelined (run_chain_pair) with the same
    results as stage by stage.  ``pipelined``: False (stage by stage), "free"
    (the chains without a stage barrier between them), "stagger" (each
    chain's products started when the other chain's pairs reach their
    eigensolvers) or "auto": staggered in mixed precision, where the
    one-workgroup eigensolvers dominate a pair's time, free in FP64, where
    the products do and two pairs alone fill the GPU less well than four."""
    from .directional_mover import run_chain_pair
    nx, ny = ipeps.nx, ipeps.ny
    if pipelined and mover.projector_calculator.projectors == "half-system":
        if pipelined == "auto":
            pipelined = "stagger" if mover.projector_calculator.precision == "mixed" else "free"
        stagger = pipelined == "stagger"
        run_chain_pair(mover, ipeps, [(xi, 0) for xi in range(nx)],
                       [((nx - xi + 1) % nx, 2) for xi in range(nx)], stagger)
        run_chain_pair(mover, ipeps, [((ny - yi + 1) % ny, 1) for yi in range(ny)],
                       [(yi, 3) for yi in range(ny)], stagger)
        return
    for xi in range(nx):
        mover.stage(ipeps, [(xi, 0), ((nx - xi + 1) % nx, 2)])
    for yi in range(ny):
        mover.stage(ipeps, [((ny - yi + 1) % ny, 1), (yi, 3)])


def stage_moves(ipeps):
    """The stages of one iteration in ctmrg_iteration's order."""
    nx, ny = ipeps.nx, ipeps.ny
    return ([[(xi, 0), ((nx - xi + 1) % nx, 2)] for xi in range(nx)] +
            [[((ny - yi + 1) % ny, 1), (yi, 3)] for yi in range(ny)])


def ctmrg_iteration_many(pairs):
    """One CTMRG iteration of each of several independent networks (a
    parameter sweep: instances of uneven D and chi), ``pairs`` a list of
    (mover, ipeps).  Stage s of every network runs in one round
    (run_stages), so one instance's single-workgroup eigensolves overlap the
    others' matrix-core work; each network sees exactly the operations of
    ctmrg_iteration, in the same order, so its result is the same."""
    from .directional_mover import run_stages
    plans = [(m, net, stage_moves(net)) for m, net in pairs]
    for s in range(max(len(st) for _, _, st in plans)):
        run_stages([(m, net, st[s]) for m, net, st in plans if s < len(st)])


def ctmrg(ipeps, config, mover=None):
    """Runs ``config.steps`` CTMRG iterations; with ``config.conv_tol`` set,
    stops early once the normalised corner spectra change by less than it.

    With ``config.graph`` set, once an iteration has run with every boundary
    dimension at chi (so its workspaces and tile choices are those of the
    fixed shapes), the next iteration is captured in a HIP graph and the
    remaining ones replay it (graph.py); a replay in which a projector pair
    kept fewer than chi singular values ends the replays, and the following
    iterations run uncaptured until the shapes have saturated again.  The
    capture is switched off for now (graph.CAPTURE_ENABLED), and
    ``config.graph`` then raises NotImplementedError before any work."""
    from . import graph as _graph
    from .graph import CapturedIteration, saturated
    if getattr(config, "graph", False) and not _graph.CAPTURE_ENABLED:
        raise NotImplementedError(_graph._DISABLED)
    mover = mover if mover is not None else DirectionalMover(config, ipeps.sketch)
    disable = config.disable_progressbar
    pbar = tqdm(desc="Renormalizing boundary tensors", unit=" sweeps", disable=disable)
    prev = corner_spectra(ipeps) if config.conv_tol else None
    use_graph = bool(getattr(config, "graph", False))
    graph, ran_saturated = None, False
    done = 0
    for _ in range(config.steps):
        pbar.update(1)
        if use_graph and graph is None and ran_saturated and saturated(ipeps):
            graph = CapturedIteration([(mover, ipeps)])
        if graph is not None:
            graph.replay()
            if not graph.counts_ok():
                graph, ran_saturated = None, False
        else:
            ran_saturated = use_graph and saturated(ipeps)
            ctmrg_iteration(mover, ipeps)
        done += 1
        if config.conv_tol:
            cur = corner_spectra(ipeps)
            if _spectra_change(prev, cur) < config.conv_tol:
                break
            prev = cur
    pbar.close()
    return done
