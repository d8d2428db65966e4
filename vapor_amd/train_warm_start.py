"""Train the learned pressure warm start on pairs the solver generates.

    python -m vapor_amd.train_warm_start [--protocol pairs64 | scenes128]
                                         [--epochs E] [--lr LR] [--batch B] [--out PATH]

Two training protocols:

- pairs64 (default): 20,000 pairs from 64^3 runs of vapor's simple
  semi-Lagrangian step (the 64^3 reference case's method, with a capped
  Gaussian emitter and no decay) over 200 seeded scenes of 100 steps each,
  every scene with its own emitter centre in [0.3,0.7]^2 x [0.1,0.2] (z up),
  radius in [0.05, 0.1] and buoyancy in [0.5, 2.0]. Weights:
  vapor_amd/weights/unet_warm_start.bin.
- scenes128: 800 pairs from 10 scenes of the 128^3 benchmark's own family
  (reflection MacCormack, seeds 1..10; the benchmark's scene is seed 0 and
  is not trained on), 40 steps each. Weights: vapor_amd/weights/unet_128.bin.

Every warm tolerance projection of those runs records (residual of its
starting pressure, the correction the solve applied), both normalised by the
residual's RMS, on the device; the U-Net then trains on them with Adam on the
library's own forward and backward kernels. The FP32 weights are written to
PATH, with the run's record beside them (PATH with a .json suffix).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent

PROTOCOLS = {
    # n, scenes, steps per scene, scene family, default epochs
    "pairs64": (64, 200, 100, "simple", 6),
    "scenes128": (128, 10, 40, "reflection", 14),
}


def scene(seed: int, mac3d):
    """The benchmark's scene family: emitter centre in [0.3,0.7]^2 x [0.1,0.2]
    (z up), radius in [0.05, 0.1], buoyancy in [0.5, 2.0]; omega 1.8, relative
    L2 tolerance 1e-5 checked every 32 sweeps, previous-pressure warm start."""
    rng = np.random.default_rng(seed)
    a, height, c = rng.uniform(0.3, 0.7), rng.uniform(0.1, 0.2), rng.uniform(0.3, 0.7)
    p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
    p.emitterCenterX, p.emitterCenterY, p.emitterCenterZ = a, c, height
    p.emitterRadius = rng.uniform(0.05, 0.1)
    p.buoyancy = rng.uniform(0.5, 2.0)
    p.sorOmega = 1.8
    p.update(solve_mode=1, tolerance=1e-5, max_sweeps=20000, check_interval=32, warm_start=1)
    return p


def simple_scene(seed: int, mac3d):
    """The 64^3 reference case's step (simple semi-Lagrangian, no reflection,
    capped-Gaussian emitter of strength 1, no decay) with the scene family's
    seeded emitter and buoyancy, solved to relative L2 1e-5 with omega 1.8 and
    the previous-pressure warm start."""
    p = scene(seed, mac3d)
    base = mac3d.MacGridParameters3D(mac3d.SimpleSemiLagrangian)
    p.update(reflection=base.reflection, maccormack_vel=base.maccormack_vel, maccormack_smoke=base.maccormack_smoke,
             smoke_decay=0.0, source_strength=1.0, emitter_profile=1)
    return p


def weights_path(n: int) -> Path:
    """Weights trained on the n^3 benchmark family (protocol scenes128 for n = 128)."""
    return PACKAGE / "weights" / f"unet_{n}.bin"


def warm_start_weights() -> Path:
    """Weights of the default protocol (pairs64)."""
    return PACKAGE / "weights" / "unet_warm_start.bin"


def train(n: int, scenes: int, steps: int, epochs: int, lr: float, batch: int = 8, seed: int = 1234,
          family: str = "reflection", log=print):
    from vapor_amd import mac3d

    make = simple_scene if family == "simple" else scene
    per_step = 1 if family == "simple" else 2
    data = mac3d.Dataset(n, scenes * steps * per_step)
    t0 = time.perf_counter()
    for s in range(1, scenes + 1):
        sim = mac3d.GpuMacSimulation3D(n, 1.0, make(s, mac3d))
        sim.record(data)
        for _ in range(steps):
            sim.advance(1.0 / 60.0)
        sim.record(None)
        sim.close()
        if s % 10 == 0 or s == scenes:
            log(f"scene {s}: {len(data)} pairs ({time.perf_counter() - t0:.1f} s)")
    pairs = len(data)
    net = mac3d.UNet(seed)
    losses = []
    for e in range(epochs):
        t1 = time.perf_counter()
        losses.append(net.train(data, 1, batch, lr, seed=e))
        log(f"epoch {e}: mean loss {losses[-1]:.5f} ({time.perf_counter() - t1:.1f} s)")
    data.close()
    return net, losses, pairs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--protocol", choices=sorted(PROTOCOLS), default="pairs64")
    ap.add_argument("--epochs", type=int, default=0, help="epochs (default: the protocol's)")
    ap.add_argument("--lr", type=float, default=4e-4)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--out", type=str, default=None)
    args = ap.parse_args()
    n, scenes, steps, family, epochs = PROTOCOLS[args.protocol]
    epochs = args.epochs or epochs
    log = lambda *a: print(*a, flush=True, file=sys.stderr)  # noqa: E731
    t0 = time.perf_counter()
    net, losses, pairs = train(n, scenes, steps, epochs, args.lr, args.batch, family=family, log=log)
    out = Path(args.out) if args.out else (warm_start_weights() if args.protocol == "pairs64" else weights_path(n))
    out.parent.mkdir(parents=True, exist_ok=True)
    net.save(out)
    record = {"weights": out.name, "protocol": args.protocol, "n": n, "scenes": scenes, "steps": steps,
              "family": family, "pairs": pairs, "epochs": epochs, "lr": args.lr, "batch": args.batch,
              "losses": [round(x, 5) for x in losses], "seconds": round(time.perf_counter() - t0, 1)}
    out.with_suffix(".json").write_text(json.dumps(record, indent=1) + "\n")
    print(json.dumps(record))


if __name__ == "__main__":
    main()
