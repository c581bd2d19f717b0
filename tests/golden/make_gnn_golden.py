"""Write golden forward vectors of the reference GNN model.

Run once in the build container (where /root/reference exists):

    python tests/golden/make_gnn_golden.py

The model forward (``build_model`` + ``stage_forward``) is pinned by no closed
form, so this script loads the reference's ``latec/levels.py`` and
``latec/gnn_predictor.py`` by file path under a stub package (the real package
``__init__`` imports ``zstandard``, which this image lacks; these two modules
need only NumPy and torch) and records, per case, into ``gnn_golden.npz``:

* the seeded model's ``state_dict`` (float32);
* the input field (B, N) in [0, 1] (float32);
* per predicted stage: the query flat indices and the head outputs
  ``mu`` and ``log_b`` of a float32 run over the whole schedule (the
  reference's ``head_of`` forces float32, so the model cannot run in float64
  end to end), with the
  true values revealed after each stage (no noise), which is
  ``scripts/train_gnn.py:run_stages`` with ``teacher_force`` and zero noise.

Nothing under ``tests/`` imports the reference; the tests read the npz only.
"""

from __future__ import annotations

import importlib.util
import os
import sys
import types

import numpy as np
import torch

REF = "/root/reference/latec"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gnn_golden.npz")

# (name, shape, levels, stride, d, agg, max_radius, batch, eb)
CASES = [
    ("c2d", (20, 24), 2, 4, 8, 2, 4, 2, 0.01),
    ("c2d_r64", (17, 19), 3, 8, 16, 1, 64, 1, 0.003),
    ("c3d", (9, 8, 10), 2, 4, 8, 2, 4, 1, 0.02),
    ("c4d", (6, 5, 6, 4), 1, 2, 8, 2, 2, 1, 0.01),
]


def load_reference():
    pkg = types.ModuleType("latec")
    pkg.__path__ = [REF]
    sys.modules["latec"] = pkg
    mods = {}
    for name in ("levels", "gnn_predictor"):
        spec = importlib.util.spec_from_file_location(f"latec.{name}", f"{REF}/{name}.py")
        mod = importlib.util.module_from_spec(spec)
        sys.modules[spec.name] = mod
        spec.loader.exec_module(mod)
        mods[name] = mod
    return mods["gnn_predictor"]


def run_case(gp, shape, levels, stride, d, agg, max_radius, batch, eb, seed):
    torch.manual_seed(seed)
    model = gp.build_model(d, agg).eval()
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    x = torch.from_numpy(rng.random((batch, n)).astype(np.float32))
    geoms, _ = gp.build_stage_geoms(shape, levels, stride, 1, max_radius, torch,
                                    None, agg)
    E = torch.zeros(batch, n, len(shape), d)
    out = {"x": x.numpy()}
    ctx = None
    with torch.no_grad():
        for i in range(1, len(geoms)):
            g0, g1 = geoms[i - 1], geoms[i]
            (mu, lb), E, ctx = gp.stage_forward(model, E, g0, g1, x[:, g0.query_idx],
                                               torch, finalize_ctx=ctx, eb=eb)
            out[f"q{i}"] = g1.query_idx.numpy()
            out[f"mu{i}"] = mu.numpy()
            out[f"lb{i}"] = lb.numpy()
    out["q0"] = geoms[0].query_idx.numpy()
    out["n_stages"] = np.int64(len(geoms))
    sd = {k: v.detach().numpy() for k, v in model.state_dict().items()}
    return out, sd


def main():
    gp = load_reference()
    arrays = {}
    for ci, (name, shape, levels, stride, d, agg, mr, batch, eb) in enumerate(CASES):
        out, sd = run_case(gp, shape, levels, stride, d, agg, mr, batch, eb, 100 + ci)
        arrays[f"{name}/meta"] = np.array([levels, stride, d, agg, mr, batch, *shape],
                                          np.int64)
        arrays[f"{name}/eb"] = np.float64(eb)
        for k, v in out.items():
            arrays[f"{name}/{k}"] = v
        for k, v in sd.items():
            arrays[f"{name}/p/{k}"] = v
        print(name, shape, "stages", int(out["n_stages"]), "params", len(sd))
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
