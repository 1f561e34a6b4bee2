"""iPEPS facade (acetn/ipeps/ipeps.py:20-137) on the MI355X engine."""
import torch
from tqdm import tqdm

from .. import _hip
from ..model.model_factory import model_factory
from ..renormalization.ctmrg import ctmrg
from ..renormalization.projectors import SketchCounter
from .ipeps_config import IpepsConfig
from .tensor_network import TensorNetwork


class Ipeps(TensorNetwork):
    def __init__(self, config, ipeps=None):
        self.config = IpepsConfig(**config)
        self.dtype = self.config.dtype
        self.device = self.config.device
        _hip.require_gpu()
        # single-GPU engine (the torchrun mode of utils/distributed.py is out of scope)
        self.rank, self.world_size, self.is_distributed = 0, 1, False
        seed = self.config.ctmrg.sketch_seed
        if seed is None:
            seed = int(torch.randint(0, 2 ** 62, (1,)).item())
        self.sketch = SketchCounter(seed)
        super().__init__(ipeps, self.config.TN, self.dtype, self.device)
        self.site_states_initialized = False

    def set_chi(self, chi):
        self.dims['chi'] = chi
        tqdm.write(f"Boundary bond dimension set: chi={chi}")

    def set_model(self, model_cls, params, name=None):
        name = "custom" if name is None else name.lower()
        model_factory.register(name, model_cls)
        self.config.model.name = name
        self.config.model.params = params

    def set_model_params(self, **kwargs):
        for key, val in kwargs.items():
            self.config.model.params[key] = val
            tqdm.write(f"Model parameter set: {key}={val}")

    def renormalize(self):
        return ctmrg(self, self.config.ctmrg)

    def evolve(self, dtau, steps, model=None):
        """ipeps.py:107-123."""
        from ..evolution.evolve import evolve
        model = model_factory.create(self.config.model) if model is None else model
        if not self.site_states_initialized:
            self.initialize_site_tensors(model.initial_site_state)
        self.renormalize()
        evolve(self, dtau, steps, model, self.config.evolution)
        self.renormalize()

    def measure(self, model=None):
        from ..measurement.measure import measure
        model = model_factory.create(self.config.model) if model is None else model
        return measure(self, model)
