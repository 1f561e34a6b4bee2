"""CPU checks of the drop-in boundary: the library loads and exports every
symbol include/acetn_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "acetn_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(acetn_hip_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "acetn_hip_half_projectors" in syms and "acetn_hip_renormalize_boundary" in syms
    assert len(syms) >= 12


def test_library_exports_every_declared_symbol():
    from acetn import _hip
    assert os.path.exists(_hip.LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(_hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_hip_backend_refuses_without_gpu():
    import torch
    from acetn import _hip
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        _hip.require_gpu()


def test_config_validation():
    import torch
    from acetn.ipeps.ipeps_config import IpepsConfig
    cfg = IpepsConfig(TN={'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': 2, 'chi': 8}},
                      evolution={'backend': 'cutensor'}, device="cpu")
    assert cfg.evolution.backend == "hip"
    assert cfg.device.type == "cuda"
    # the reference's default backend name runs on the engine too
    cfg = IpepsConfig(TN={'nx': 2, 'ny': 2}, evolution={'backend': 'torch'})
    assert cfg.evolution.backend == "hip"
    with pytest.raises(ValueError):
        IpepsConfig(TN={'nx': 2, 'ny': 2}, evolution={'backend': 'numpy'})
    with pytest.raises(TypeError):
        IpepsConfig(TN={'nx': 2, 'ny': 2}, ctmrg={'no_such_key': 1})
    assert IpepsConfig(TN={'nx': 2, 'ny': 2}, dtype="float32").dtype == torch.float64
    assert IpepsConfig(TN={'nx': 2, 'ny': 2}, dtype=torch.float32).dtype == torch.float64
    with pytest.raises(NotImplementedError, match="real"):
        IpepsConfig(TN={'nx': 2, 'ny': 2}, dtype="complex128")


def test_ctmrg_graph_config_raises_while_switched_off():
    """config.graph=True fails before any work while the capture is off
    (runs without a GPU: the check precedes every device call)."""
    from acetn.renormalization.ctmrg import ctmrg

    class C:
        graph = True
    with pytest.raises(NotImplementedError, match="switched off"):
        ctmrg(None, C())


def test_learned_start_network_save_load_roundtrip(tmp_path):
    """MLP.save writes plain tensors that load with weights_only=True; the
    layer sizes are checked (host-side, no kernels)."""
    import torch
    from acetn.evolution.learned_start import MLP
    m = MLP(40, 12, hidden=16, seed=3, device="cpu")
    p = tmp_path / "net.pt"
    m.save(str(p))
    st = torch.load(str(p), weights_only=True)
    assert st["fin"] == 40 and st["hidden"] == 16 and st["fout"] == 12
    m2 = MLP.load(str(p), device="cpu")
    assert torch.equal(m2.params, m.params)
    st["params"] = st["params"][:-1]
    torch.save(st, str(p))
    with pytest.raises(ValueError, match="parameter count"):
        MLP.load(str(p), device="cpu")


def test_boundary_refuses_host_and_non_float64_tensors():
    """The kernels read FP64 device memory: a float32 (or complex) site
    tensor is refused at the boundary instead of being read as FP64 bytes,
    and a host tensor instead of being handed to a kernel as a device
    address (both guards run before any device call)."""
    import torch
    from acetn import _hip
    A = torch.zeros(2, 2, 2, 2, 2, dtype=torch.float64)
    C = [torch.zeros(4, 4, dtype=torch.float64) for _ in range(4)]
    E = [torch.zeros(4, 4, 2, 2, dtype=torch.float64) for _ in range(4)]
    for bad in (torch.float32, torch.complex128):
        with pytest.raises(TypeError, match="float64"):
            _hip.make_site(A.to(bad), C, E)
        with pytest.raises(TypeError, match="float64"):
            _hip.make_site(A, C[:3] + [C[3].to(bad)], E)
    with pytest.raises(ValueError, match="GPU"):
        _hip.make_site(A, C, E)
    with pytest.raises(ValueError, match="GPU"):
        _hip.ptr(A)
    assert _hip.ptr(None).value is None
