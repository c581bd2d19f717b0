"""latec_amd -- MI355X-native LATEC codec hot path.

The reference's public API (latec/__init__.py:4-8): ``compress``,
``decompress`` and ``InterpPredictor`` for the plain interpolation codec,
backed by hand-written gfx950 HIP kernels in ``liblatec.so``.
"""

from .codec import compress, decompress
from .predictor import InterpPredictor

__version__ = "0.1.0"

__all__ = ["compress", "decompress", "InterpPredictor", "__version__"]
