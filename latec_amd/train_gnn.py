"""Train a GNN checkpoint on the GPU (the reference's scripts/train_gnn.py role).

    python -m latec_amd.train_gnn --out ckpt.pt --steps 400

    python -m latec_amd.train_gnn --out ckpt.pt --steps 400 --c5   # the C5 setup

Synthetic batches of 2-D and 4-D fields (Gaussian random fields with random
correlation lengths, power-law turbulence, hyperplane jumps) are passed
through the codec's own progressive schedule, with teacher forcing or (with
--closed-loop) the codec's own quantized reconstruction as each stage's
context, at bounds drawn log-uniformly from --eb-lo..--eb-hi; the loss is the
discretized-Laplace code length in bits.  Forward and backward run the
liblatec trainer kernels (--precision bf16: BF16 MFMA GEMMs); the optimizer
is AdamW on FP32 master weights.  --c5 selects batches of 32 fields (16 x
256^2, 16 x 16^4), eb 1e-3, closed loop and BF16.  The output is a v7
checkpoint.
"""

from __future__ import annotations

import argparse
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--d", type=int, default=64)
    ap.add_argument("--agg", type=int, default=2)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--batch2d", type=int, default=16)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--batch4d", type=int, default=4)
    ap.add_argument("--edge4d", type=int, default=16)
    ap.add_argument("--levels", type=int, default=3)
    ap.add_argument("--eb-lo", type=float, default=1e-4)
    ap.add_argument("--eb-hi", type=float, default=1e-2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from this checkpoint")
    ap.add_argument("--closed-loop", action="store_true")
    ap.add_argument("--precision", choices=("fp32", "bf16"), default="fp32")
    ap.add_argument("--c5", action="store_true", help="the C5 batch, bound and targets")
    args = ap.parse_args(argv)
    from . import gnn_model as GM
    from .gnn_train import C5, GnnTrainer, train

    params = None
    d, agg = args.d, args.agg
    if args.init:
        d, agg, params = GM.load_checkpoint(args.init)
    tr = GnnTrainer(d, agg, seed=args.seed, params=params,
                    precision="bf16" if args.c5 else args.precision)
    t0 = time.time()

    def log(msg):
        print(f"[train {time.time() - t0:7.1f}s] {msg}", file=sys.stderr, flush=True)

    if args.c5:
        cfg = dict(C5)
    else:
        cfg = dict(batch2d=args.batch2d, shape2d=(args.crop, args.crop), batch4d=args.batch4d,
                   shape4d=(args.edge4d,) * 4, levels=args.levels,
                   eb_range=(args.eb_lo, args.eb_hi), closed_loop=args.closed_loop)
    hist = train(tr, args.steps, lr=args.lr, warmup=args.warmup, seed=args.seed, log=log, **cfg)
    tr.save(args.out, extra={"train_steps": args.steps, "final_bits": float(hist[-1])})
    log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
