"""acetn on AMD Instinct MI355X (gfx950).

Keeps ace-tn's public API (``from acetn.ipeps import Ipeps``; ``Ipeps(config)``,
``.evolve()``, ``.measure()``, ``.renormalize()``) and runs its hot path --
the CTMRG step, the full-update bond solve and the reduced density
matrices -- in hand-written HIP kernels behind the C-ABI of
include/acetn_hip.h.
"""
__version__ = "0.1.3+mi355x"
