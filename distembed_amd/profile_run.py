r"""The R-MAT scale-20, d = 128 workload of one gather form alone, for a
profiler to watch: one context at the tiling kept under profiles/, warm-up
epochs, then timed epochs replayed from the captured epoch graph. No event
timing of its own (the profiler times the kernels).

  python -m distembed_amd.profile_run fp32|bf16 [--epochs 100] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

from distembed_amd import api

ROOT = Path(__file__).resolve().parent.parent


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("gather", choices=["fp32", "bf16"])
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    a = ap.parse_args()
    B, ns, lr = 16384, 5, 0.02
    g = api.Graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=0)
    X0 = api.init_uniform(g.n, 128, -0.5, 0.5, seed=0)
    kept = json.loads((ROOT / "profiles" / f"force_kernel_rmat20_d128_{a.gather}_tuning.json").read_text())
    ctx = api.ForceEmbed(g.n, 128, gather=api.BF16 if a.gather == "bf16" else api.FP32)
    try:
        ctx.upload_csr(g)
        ctx.upload_x(X0)
        ctx.configure(B, lr, 5.0)
        ctx.set_tuning(**kept["tuning"])
        ctx.set_batch_tunings([tuple(t) for t in kept["per_batch"]])
        ctx.train(a.warmup, B, ns, lr)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.train(a.epochs, B, ns, lr)
        ctx.synchronize()
        dt = time.perf_counter() - t0
    finally:
        ctx.close()
    print(f"profile_run {a.gather}: {a.epochs} epochs in {dt:.3f} s "
          f"({dt / a.epochs * 1e3:.3f} ms per epoch, host clock)", flush=True)


if __name__ == "__main__":
    main()
