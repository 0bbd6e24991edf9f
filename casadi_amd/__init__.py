"""casadi_amd: an MI355X-native batched evaluator for CasADi SX expression graphs.

Keeps CasADi's Function / Function.map(N, parallelization) / Linsol API and
adds the ``"hip"`` parallelization, whose hot path is a set of hand-written
gfx950 kernels behind the C ABI in ``include/casadi_hip.h``.
"""
from ._lib import CasadiHipError  # noqa: F401  (fails loudly without the native library)
from .ops import OP  # noqa: F401
from .sx import *  # noqa: F401,F403
from .sx import Sparsity, SX  # noqa: F401
from .function import Function  # noqa: F401
from .map import MapFunction  # noqa: F401

__version__ = "0.1.0"
