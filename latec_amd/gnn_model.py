"""GNN checkpoints: parameter names, loading, saving, initialization.

The checkpoint format is the reference's v7 (``latec/gnn_predictor.py:58,
945-981``): a ``torch.save`` dict ``{"version": 7, "d", "agg_level",
"state_dict"}`` whose ``state_dict`` keys are the reference model's module
paths, read here with ``torch.load(weights_only=True)``.  The hash stored in
streams is the first 16 bytes of the file's sha256 (``file_sha256``,
gnn_predictor.py:936-942).
"""

from __future__ import annotations

import hashlib
import math
from pathlib import Path

import numpy as np

CKPT_VERSION = 7


def param_shapes(d: int, agg: int) -> list[tuple[str, tuple[int, ...]]]:
    """(name, shape) of every parameter, in the packed order the C-ABI takes
    (module order of ``build_model``, gnn_predictor.py:593-933)."""
    h, ds = 2 * d, max(1, d // 4)
    de, ns = d + ds, 2 * agg + 1
    out = [
        ("init.net.0.weight", (d, 1)), ("init.net.0.bias", (d,)),
        ("init.net.2.weight", (d, d)), ("init.net.2.bias", (d,)),
        ("dir_state.table", (ns, ds)),
        ("dir.net.0.weight", (h, de + 2)), ("dir.net.0.bias", (h,)),
        ("dir.net.2.weight", (d, h)), ("dir.net.2.bias", (d,)),
        ("bidir.net.0.weight", (h, 2 * de + 2)), ("bidir.net.0.bias", (h,)),
        ("bidir.net.2.weight", (d, h)), ("bidir.net.2.bias", (d,)),
    ]
    for pool in ("line_pool", "axis_pool"):
        out += [
            (f"{pool}.wk.0.weight", (d, d)), (f"{pool}.wk.0.bias", (d,)),
            (f"{pool}.wk.2.weight", (1, d)), (f"{pool}.wk.2.bias", (1,)),
            (f"{pool}.wv.weight", (d, d)), (f"{pool}.wv.bias", (d,)),
            (f"{pool}.null_k", ()), (f"{pool}.null_v", (d,)),
        ]
    out += [
        ("head.net.0.weight", (h, d + 1)), ("head.net.0.bias", (h,)),
        ("head.net.2.weight", (2, h)), ("head.net.2.bias", (2,)),
        ("mix.net.0.weight", (h, 2 * d)), ("mix.net.0.bias", (h,)),
        ("mix.net.2.weight", (d, h)), ("mix.net.2.bias", (d,)),
    ]
    return out


def file_sha256(path) -> bytes:
    digest = hashlib.sha256()
    with open(path, "rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            digest.update(block)
    return digest.digest()


def load_checkpoint(path):
    """(d, agg_level, {name: float32 ndarray}) of a v7 checkpoint; the
    reference's messages on a wrong version or a missing aggregation level."""
    import torch

    ckpt = torch.load(Path(path), map_location="cpu", weights_only=True)
    version = int(ckpt.get("version", 1))
    if version != CKPT_VERSION:
        raise ValueError(
            "Axial GNN checkpoint format v7 is required. Retrain with "
            "scripts/train_gnn.py."
        )
    d = int(ckpt["d"])
    if "agg_level" not in ckpt:
        raise ValueError(
            "checkpoint is missing 'agg_level'; retrain with scripts/train_gnn.py."
        )
    agg = int(ckpt["agg_level"])
    sd = ckpt["state_dict"]
    params = {}
    for name, shape in param_shapes(d, agg):
        if name not in sd:
            raise ValueError(f"checkpoint state_dict lacks {name}")
        t = sd[name].detach().to("cpu").float().numpy()
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")
        params[name] = np.array(t, dtype=np.float32, copy=True).reshape(shape)
    return d, agg, params


def save_checkpoint(path, d: int, agg: int, params: dict, extra: dict | None = None) -> None:
    """Write a v7 checkpoint the reference's loader also accepts."""
    import torch

    sd = {name: torch.from_numpy(np.array(params[name], np.float32).reshape(shape))
          for name, shape in param_shapes(d, agg)}
    ckpt = {"version": CKPT_VERSION, "d": int(d), "agg_level": int(agg), "state_dict": sd}
    if extra:
        ckpt.update(extra)
    torch.save(ckpt, Path(path))


def pack(params: dict, d: int, agg: int) -> np.ndarray:
    """The packed float32 block of ``latec_gnn_create``."""
    return np.concatenate([np.asarray(params[n], np.float32).reshape(-1)
                           for n, _ in param_shapes(d, agg)])


def init_params(d: int, agg: int, seed: int = 0) -> dict:
    """Fresh weights: uniform(+-1/sqrt(fan_in)) for Linear layers (PyTorch's
    default scale), DirState rows ~ N(0, 1/ds), null keys ~ N(0, 1/d), null
    values 0, and the head's log-scale bias at 8 (an unconfident start,
    gnn_predictor.py:749-758)."""
    rng = np.random.default_rng(seed)
    out = {}
    ds = max(1, d // 4)
    for name, shape in param_shapes(d, agg):
        if name == "dir_state.table":
            v = rng.standard_normal(shape) * ds ** -0.5
        elif name.endswith("null_k"):
            v = rng.standard_normal(shape) * d ** -0.5
        elif name.endswith("null_v"):
            v = np.zeros(shape)
        else:
            wname = name.rsplit(".", 1)[0] + ".weight"
            fan_in = dict(param_shapes(d, agg))[wname][1]
            bound = 1.0 / math.sqrt(fan_in)
            v = rng.uniform(-bound, bound, shape)
        out[name] = np.asarray(v, np.float32)
    out["head.net.2.bias"][1] = 8.0
    return out
