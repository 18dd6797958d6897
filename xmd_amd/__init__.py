"""MI355X engine for xmd's per-step force-and-integrate loop.

Host side: YAML front end (params), synthetic systems (synth) and the
ctypes driver (engine) of the C-ABI library libxmdhip.so, whose HIP kernels
do all the compute.
"""
from . import _abi  # noqa: F401
from ._abi import TERM, TERMS, ALL_TERMS, LIST, PREC_FP64, PREC_MIXED  # noqa: F401

__version__ = "0.1.0"
