"""Train the learned pressure warm start on pairs the solver generates.

    python -m vapor_amd.train_warm_start [--n 128] [--scenes 10] [--steps 40]
                                         [--epochs 14] [--lr 4e-4] [--out PATH]

Each training scene is the benchmark's reflection MacCormack workload with
its own seeded emitter (scene seeds 1..scenes; the benchmark's own scene is
seed 0 and is not trained on). Every warm tolerance projection of those
runs records (residual of its starting pressure, the correction the solve
applied), both normalised by the residual's RMS, on the device; the U-Net
then trains on them with Adam on the library's own forward and backward
kernels. The FP32 weights are written to PATH (default
vapor_amd/weights/unet_<n>.bin).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent


def scene(seed: int, mac3d):
    """The benchmark's scene family: emitter centre in [0.3,0.7]^2 x [0.1,0.2]
    (z up), radius in [0.05, 0.1], buoyancy in [0.5, 2.0]; omega 1.8, relative
    L2 tolerance 1e-5, previous-pressure warm start."""
    rng = np.random.default_rng(seed)
    a, height, c = rng.uniform(0.3, 0.7), rng.uniform(0.1, 0.2), rng.uniform(0.3, 0.7)
    p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
    p.emitterCenterX, p.emitterCenterY, p.emitterCenterZ = a, c, height
    p.emitterRadius = rng.uniform(0.05, 0.1)
    p.buoyancy = rng.uniform(0.5, 2.0)
    p.sorOmega = 1.8
    p.update(solve_mode=1, tolerance=1e-5, max_sweeps=20000, check_interval=16, warm_start=1)
    return p


def weights_path(n: int) -> Path:
    return PACKAGE / "weights" / f"unet_{n}.bin"


def train(n: int, scenes: int, steps: int, epochs: int, lr: float, batch: int = 8, seed: int = 1234,
          log=print):
    from vapor_amd import mac3d

    data = mac3d.Dataset(n, scenes * steps * 2)
    t0 = time.perf_counter()
    for s in range(1, scenes + 1):
        sim = mac3d.GpuMacSimulation3D(n, 1.0, scene(s, mac3d))
        sim.record(data)
        for _ in range(steps):
            sim.advance(1.0 / 60.0)
        sim.record(None)
        sim.close()
        log(f"scene {s}: {len(data)} pairs ({time.perf_counter() - t0:.1f} s)")
    net = mac3d.UNet(seed)
    losses = []
    for e in range(epochs):
        t1 = time.perf_counter()
        losses.append(net.train(data, 1, batch, lr, seed=e))
        log(f"epoch {e}: mean loss {losses[-1]:.5f} ({time.perf_counter() - t1:.1f} s)")
    data.close()
    return net, losses


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--scenes", type=int, default=10)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=14)
    ap.add_argument("--lr", type=float, default=4e-4)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--out", type=str, default=None)
    args = ap.parse_args()
    log = lambda *a: print(*a, flush=True, file=sys.stderr)  # noqa: E731
    net, losses = train(args.n, args.scenes, args.steps, args.epochs, args.lr, args.batch, log=log)
    out = Path(args.out) if args.out else weights_path(args.n)
    out.parent.mkdir(parents=True, exist_ok=True)
    net.save(out)
    print(json.dumps({"weights": str(out), "n": args.n, "pairs": args.scenes * args.steps * 2, "epochs": args.epochs,
                      "losses": [round(x, 5) for x in losses]}))


if __name__ == "__main__":
    main()
