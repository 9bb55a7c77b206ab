"""Trains the JKR surrogate on the GPU and writes the weight file.

python -m bfm.train_surrogate [out.npz]   (run with biofilm-abms_amd on sys.path)
1e6 samples of the exact FP64 solve (overlap U[-0.2 R, 0.3 R], gamma
U[10, 1000] Pa um, seed 42), batch 4096, Adam, cosine-decayed lr.
"""
import sys
import time

import numpy as np

import bfm
from bfm import surrogate as sur


def main(out=sur.DEFAULT_WEIGHTS, n=1_000_000, epochs=80, batch=4096, lr=1e-3, seed=42):
    ctx = bfm.Context(1024, params=bfm.core.make_params(jkr_mode=bfm.core.JKR_SURROGATE))
    t0 = time.time()
    params, losses = sur.train(ctx, n_samples=n, epochs=epochs, batch=batch, lr=lr, seed=seed)
    dt = time.time() - t0
    print(f"trained in {dt:.1f} s; loss per epoch: {np.array2string(losses, precision=3)}",
          flush=True)
    sur.save(out, params, losses, n_samples=n, epochs=epochs, batch=batch, lr=lr, seed=seed,
             train_seconds=dt)
    print(f"wrote {out}")


if __name__ == "__main__":
    main(*sys.argv[1:2])
