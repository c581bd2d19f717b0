"""latec_amd -- MI355X-native LATEC codec hot path.

The reference's public API (latec/__init__.py:4-8): ``compress`` and
``decompress`` for the plain codec with any predictor (``InterpPredictor``
runs as fused gfx950 stage kernels in ``liblatec.so``), and the learned
codec ``GNNCodec`` (``GNNCompressorCodec``).  The GNN codec module imports
torch on first use, so ``import latec_amd`` stays light.
"""

from .codec import compress, decompress
from .predictor import InterpPredictor

__version__ = "0.1.0"

__all__ = ["compress", "decompress", "InterpPredictor", "GNNCodec", "GNNCompressorCodec",
           "__version__"]


def __getattr__(name):
    if name in ("GNNCodec", "GNNCompressorCodec"):
        from . import gnn_codec

        return getattr(gnn_codec, name)
    raise AttributeError(f"module 'latec_amd' has no attribute {name!r}")
