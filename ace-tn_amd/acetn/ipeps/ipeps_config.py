"""Configuration dataclasses (acetn/ipeps/ipeps_config.py:9-109).

Same keys and defaults as the reference, plus the engine's own:
  ctmrg.backend     -- "hip" (the only engine here; the reference's backend
                       names "torch" and "cutensor" map to it with a note)
  ctmrg.precision   -- "fp64" (plain path) | "mixed" (split-bf16 matrix-core
                       contractions with FP32 accumulation, FP64 QR/SVD)
  ctmrg.sketch_seed -- seed of the Philox range-finder sketch (None: drawn
                       once from torch's global generator, so
                       torch.manual_seed makes runs reproducible)
  ctmrg.conv_tol    -- optional early stop on the corner-spectrum change
  ctmrg.graph       -- capture one iteration in a HIP graph once the boundary
                       dimension has saturated at chi, and replay it (switched
                       off for now: True raises NotImplementedError)
  evolution.als_start -- "plain" (SVD start, als_solver.py:117-146) or
                       "learned" (MLP proposal, judged against the plain start)
  evolution.als_start_model -- path of the saved network for "learned"
Unknown keys raise TypeError, as the reference's dataclass(**kwargs) does.
"""
import os
from dataclasses import dataclass, field
from typing import Dict, Literal, Optional, Union

import torch
from tqdm import tqdm


@dataclass
class TNConfig:
    nx: int = 2
    ny: int = 2
    dims: Dict[str, int] = field(default_factory=lambda: {'phys': 2, 'bond': 2, 'chi': 20})


@dataclass
class CTMRGConfig:
    steps: int = 40
    projectors: Literal["half-system", "full-system"] = "half-system"
    svd_type: Literal["rsvd", "full-rank"] = "rsvd"
    svd_cutoff: float = 1e-12
    rsvd_niter: int = 2
    rsvd_oversampling: int = 2
    disable_progressbar: bool = False
    backend: Literal["hip"] = "hip"
    precision: Literal["fp64", "mixed"] = "fp64"
    sketch_seed: Optional[int] = None
    conv_tol: Optional[float] = None
    graph: bool = False


@dataclass
class EvolutionConfig:
    backend: Literal["hip"] = "hip"
    update_type: Literal["full", "simple"] = "full"
    use_gauge_fix: bool = True
    gauge_fix_atol: float = 1e-12
    positive_approx_cutoff: float = 1e-12
    als_niter: int = 100
    als_tol: float = 1e-15
    als_method: Literal["cholesky", "pinv"] = "cholesky"
    als_epsilon: float = 1e-12
    disable_progressbar: bool = False
    als_start: Literal["plain", "learned"] = "plain"
    als_start_model: Optional[str] = None


@dataclass
class ModelConfig:
    dtype: torch.dtype
    device: torch.device
    dim: int
    name: str = None
    params: Dict[str, float] = None


@dataclass
class IpepsConfig:
    """Configuration class for iPEPS setup."""
    TN: Dict
    dtype: Union[str, torch.dtype] = "float64"
    device: Union[str, torch.device] = "cuda"
    ctmrg: Dict = field(default_factory=lambda: {})
    evolution: Dict = field(default_factory=lambda: {})
    model: Dict = field(default_factory=lambda: {})

    def __post_init__(self):
        self.set_dtype(self.dtype)
        self.set_device(self.device)
        self.TN = TNConfig(**self.TN)
        self.ctmrg = CTMRGConfig(**self.ctmrg)
        self.evolution = EvolutionConfig(**self.evolution)
        self.validate_backend()
        self.model = ModelConfig(dtype=self.dtype, device=self.device,
                                 dim=self.TN.dims['phys'], **self.model)
        self.disable_pbar_if_slurm()

    def disable_pbar_if_slurm(self):
        """ipeps_config.py:71-81."""
        if 'SLURM_JOB_ID' in os.environ:
            job_name = os.getenv('SLURM_JOB_NAME', '')
            job_partition = os.getenv('SLURM_JOB_PARTITION', '')
            interactive = 'interactive' in job_name or job_partition == 'interactive'
            self.ctmrg.disable_progressbar = not interactive
            self.evolution.disable_progressbar = not interactive

    def set_dtype(self, dtype):
        """ipeps_config.py:86-90.  The engine stores and computes in float64:
        a real dtype of lower precision (the reference's float32) is noted
        and raised to float64 (lower precision is chosen per stage, with
        ctmrg.precision); complex dtypes are not served."""
        d = torch.float64 if dtype is None else dtype
        if isinstance(d, str):
            d = getattr(torch, d)
        if d.is_complex:
            raise NotImplementedError(
                f"dtype={d}: the MI355X engine serves real networks only (float64)")
        if d != torch.float64:
            tqdm.write(f"acetn (MI355X engine): dtype={d} computed and stored as float64; "
                       "lower precision is selected per stage with ctmrg.precision")
            d = torch.float64
        self.dtype = d

    def set_device(self, device):
        """The engine keeps every tensor on the GPU.  A 'cpu' request (the
        reference's default) is noted and replaced by the current HIP device."""
        if device is None or (isinstance(device, str) and device.startswith("cpu")) or \
                (isinstance(device, torch.device) and device.type == "cpu"):
            if device is not None:
                tqdm.write("acetn (MI355X engine): tensors live on the GPU; device='cpu' replaced "
                           "by 'cuda'")
            device = "cuda"
        self.device = torch.device(device) if isinstance(device, str) else device
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", 0)

    def validate_backend(self):
        """ipeps_config.py:103-109, re-targeted: the reference falls back from
        'cutensor' to 'torch' with a note when the extension is missing; here
        the engine's own 'hip' backend runs every request, so both of the
        reference's backend names ('torch', its default, and 'cutensor') map
        to 'hip' with a note, and other names raise."""
        for sec in (self.evolution, self.ctmrg):
            if sec.backend in ("cutensor", "torch"):
                tqdm.write(f"backend={sec.backend!r} requested; the MI355X engine runs it as "
                           "backend='hip'")
                sec.backend = "hip"
            if sec.backend != "hip":
                raise ValueError(f"backend={sec.backend!r} is not provided by the MI355X engine "
                                 "(use backend='hip')")
