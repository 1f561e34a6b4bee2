"""The tile-shape autotuner of the matrix-core products: tuned products equal
the threshold-chosen ones to rounding, the table persists through save/load,
'table' mode tunes nothing new, and a CTMRG iteration run with tuning matches
the oracle as the untuned one does."""
import copy
import os

import numpy as np
import pytest
import torch

from oracle import ctmrg as oc
from oracle import measure as om
from oracle import model as omod
from oracle.rng import SketchStream
from test_hip_ctmrg import Cfg, DevState, MixedCfg, assert_state_close, dev

pytestmark = pytest.mark.gpu


@pytest.fixture()
def hip():
    from acetn import _hip
    _hip.require_gpu()
    _hip.autotune("off")
    _hip.autotune_clear()
    yield _hip
    _hip.autotune("off")
    _hip.autotune_clear()


SHAPES = [(5184, 146, 5184, False, False), (146, 146, 5184, True, False), (1024, 66, 1024, False, False),
          (300, 200, 700, False, True), (2000, 2000, 144, False, False)]


def test_tuned_products_match_threshold_choice(hip, tmp_path):
    rng = np.random.default_rng(3)
    ref = {}
    for pieces in (0, 2, 3):
        for M, N, K, ta, tb in SHAPES:
            A = rng.standard_normal((K, M) if ta else (M, K))
            B = rng.standard_normal((N, K) if tb else (K, N))
            ref[(pieces, M, N, K)] = (A, B, hip.gemm(dev(A), dev(B), ta=ta, tb=tb, pieces=pieces).cpu().numpy())
    hip.autotune("tune")
    for (pieces, M, N, K), (A, B, C0) in ref.items():
        ta, tb = [(s[3], s[4]) for s in SHAPES if s[:3] == (M, N, K)][0]
        C1 = hip.gemm(dev(A), dev(B), ta=ta, tb=tb, pieces=pieces).cpu().numpy()
        exact = (A.T if ta else A) @ (B.T if tb else B)
        tol = {0: 1e-12, 2: 1e-4, 3: 1e-5}[pieces] * np.abs(exact).max()
        np.testing.assert_allclose(C1, exact, rtol=0, atol=tol)
        np.testing.assert_allclose(C1, C0, rtol=0, atol=tol)
    n = hip.autotune_entries()
    assert n >= len(SHAPES)
    path = os.path.join(tmp_path, "tiles.txt")
    hip.autotune_save(path)
    hip.autotune_clear()
    assert hip.autotune_entries() == 0
    assert hip.autotune_load(path) == n
    hip.autotune("table")
    A = rng.standard_normal((333, 77))
    hip.gemm(dev(A), dev(A.T.copy()))
    assert hip.autotune_entries() == n   # nothing new tuned in 'table' mode


@pytest.mark.parametrize("cfg", [Cfg, MixedCfg])
def test_ctmrg_iteration_with_tuning_matches_oracle(hip, cfg):
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi, seed = 3, 27, 91
    st0 = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=9, bias=0.5))
    ost = copy.deepcopy(st0)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    hip.autotune("tune")
    hst = DevState(st0, 2, 2, D, chi, seed)
    mover = DirectionalMover(cfg(), hst.sketch)
    for _ in range(2):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(mover, hst)
    assert hip.autotune_entries() > 0
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(hst.to_numpy(), 2, 2, omod.Heisenberg(1.0))['Energy']
    if cfg is Cfg:
        assert_state_close(hst.to_numpy(), ost, rtol=1e-9)
        assert e_h == pytest.approx(e_o, rel=1e-10)
    else:
        assert abs(e_h - e_o) <= 1e-6


@pytest.mark.parametrize("wm,var", [(2, 1), (2, 2), (2, 3), (1, 3)])
def test_wide_fp64_variants_are_exact(hip, tmp_path, wm, var):
    """The eight-wave and k-tile-32 FP64 kernels, forced through a loaded
    table entry (kind M N K flags wm wn ks ms variant), on a single-stride
    range-finder shape with and without split-K: FP64-exact against numpy;
    an entry naming a variant for a strided product is ignored on load."""
    rng = np.random.default_rng(var)
    M, N, K = 1040, 146, 1100
    A, B = rng.standard_normal((M, K)), rng.standard_normal((K, N))
    exact = A @ B
    for ks in (1, 5):
        path = os.path.join(tmp_path, f"t{ks}.txt")
        with open(path, "w") as f:
            f.write(f"0 {M} {N} {K} 3 {wm} 5 {ks} 0.0 {var}\n")
        hip.autotune_clear()
        assert hip.autotune_load(path) == 1
        hip.autotune("table")
        C = hip.gemm(dev(A), dev(B)).cpu().numpy()
        np.testing.assert_allclose(C, exact, rtol=0, atol=1e-12 * np.abs(exact).max())
    with open(path, "w") as f:
        f.write(f"0 {M} {N} {K} 2 {wm} 5 1 0.0 {var}\n")   # flags 2: not single-stride
    hip.autotune_clear()
    assert hip.autotune_load(path) == 0


@pytest.mark.parametrize("wm", [1, 2])
def test_deep_k_tile_split_variant(hip, tmp_path, wm):
    """The two-piece split product with 64-deep k-tiles (variant 1), forced
    through a loaded table entry on a planar shape whose padded K is a
    multiple of 64: with one k part it accumulates in the order of the
    32-deep kernel, so the two agree bit for bit; with four parts the part
    boundaries differ (288 against 320 columns), so they agree to the split
    products' rounding, and both match numpy to that tolerance."""
    rng = np.random.default_rng(wm)
    M, N, K = 1040, 146, 1152
    A, B = rng.standard_normal((M, K)), rng.standard_normal((K, N))
    exact = A @ B
    tol = 1e-4 * np.abs(exact).max()
    for ks in (1, 4):
        out = {}
        for var in (0, 1):
            path = os.path.join(tmp_path, f"s{ks}{var}.txt")
            with open(path, "w") as f:
                f.write(f"2 {M} {N} {K} 1 {wm} 5 {ks} 0.0 {var}\n")
            hip.autotune_clear()
            assert hip.autotune_load(path) == 1
            hip.autotune("table")
            out[var] = hip.gemm(dev(A), dev(B), pieces=2).cpu().numpy()
            np.testing.assert_allclose(out[var], exact, rtol=0, atol=tol)
        if ks == 1:
            assert np.array_equal(out[0], out[1])
        else:
            np.testing.assert_allclose(out[1], out[0], rtol=0, atol=tol)
