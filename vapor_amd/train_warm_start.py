"""Train the learned pressure warm start on pairs the solver generates.

    python -m vapor_amd.train_warm_start [--protocol coarse32 | pairs64 | scenes128]
                                         [--epochs E] [--lr LR] [--batch B] [--out PATH] [--init PATH]

Protocols:

- coarse32 (default): the 20,000-pair protocol of the 64^3 reference case's
  method (vapor's simple semi-Lagrangian step, capped Gaussian emitter, no
  decay) over 200 seeded scenes of 100 steps, every scene with its own
  emitter centre in [0.3,0.7]^2 x [0.1,0.2] (z up), radius in [0.05, 0.1] and
  buoyancy in [0.5, 2.0], plus 6,000 pairs of the 128^3 benchmark's scene
  family (reflection MacCormack, seeds 1..60, 50 steps; the benchmark's scene
  is seed 0), each pair restricted to 32^3 by cell means and read after 32
  plain sweeps from the previous pressure. (coarse32_64 and coarse32_128 use
  one source each.)
  The network then runs on 32^3 inside any solve whose grid is a multiple of
  32 (the residual restricted, the prediction prolonged back): its four
  levels see the whole domain, and it predicts exactly the smooth part of
  the correction that red-black SOR removes slowly. Held-out pairs (other
  seeds at 64^3, and the 128^3 benchmark family, restricted by 4) give the
  validation error after every epoch. Weights: vapor_amd/weights/unet_coarse32.bin.
- pairs64: the same 20,000 pairs at full 64^3 resolution
  (vapor_amd/weights/unet_warm_start.bin).
- scenes128: 800 pairs from 10 scenes of the 128^3 benchmark's own family
  (reflection MacCormack, seeds 1..10), full resolution
  (vapor_amd/weights/unet_128.bin).

Every warm tolerance projection of those runs records (residual of its
starting pressure, the correction the solve applied), both normalised by the
residual's RMS, on the device; the U-Net then trains on them with Adam on the
library's own forward and backward kernels. The FP32 weights are written to
PATH, with the run's record beside them (PATH with a .json suffix): data,
per-epoch training loss and validation error.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent

# protocol -> (sources, default epochs, network grid (0: the sources' n));
# a source is (n, first seed, scenes, steps per scene, scene family)
PROTOCOLS = {
    "coarse32": ([(64, 1, 200, 100, "simple"), (128, 1, 60, 50, "reflection")], 12, 32),
    "coarse32_64": ([(64, 1, 200, 100, "simple")], 12, 32),
    "coarse32_128": ([(128, 1, 60, 50, "reflection")], 12, 32),
    "pairs64": ([(64, 1, 200, 100, "simple")], 6, 0),
    "scenes128": ([(128, 1, 10, 40, "reflection")], 14, 0),
}
# held-out validation pairs: (n, first seed, scenes, steps, family)
VALIDATION = [(64, 5001, 4, 25, "simple"), (128, 6001, 2, 10, "reflection")]


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
    """Weights of the full-resolution 64^3 protocol (pairs64)."""
    return PACKAGE / "weights" / "unet_warm_start.bin"


def coarse_weights() -> Path:
    """Weights of the default protocol (coarse32): the network on 32^3."""
    return PACKAGE / "weights" / "unet_coarse32.bin"


def load_network(mac3d, path: Path):
    """A UNet with the weights at path and the grid its record names."""
    net = mac3d.UNet(0)
    net.load(path)
    record = path.with_suffix(".json")
    if record.exists():
        net.set_grid(int(json.loads(record.read_text()).get("grid", 0)))
    return net


def record_pairs(mac3d, data, n, seeds, steps, family, log=None):
    """Run one scene per seed at n^3 for `steps` steps, recording into data."""
    make = simple_scene if family == "simple" else scene
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        sim = mac3d.GpuMacSimulation3D(n, 1.0, make(s, mac3d))
        sim.record(data)
        for _ in range(steps):
            sim.advance(1.0 / 60.0)
        sim.record(None)
        sim.close()
        if log and ((k + 1) % 20 == 0 or k + 1 == len(seeds)):
            log(f"{n}^3 {family}: {k + 1} scenes, {len(data)} pairs ({time.perf_counter() - t0:.1f} s)")


def validation_error(net, data):
    """Relative L2 error of the network's prediction on every pair of data,
    raw and with the best scalar factor (the solver's line search fixes the
    scale): means over pairs."""
    m = data.n
    raw, scaled = [], []
    for i in range(len(data)):
        x, y = data.pair(i)
        pred = net.predict(m, x).astype(np.float64)
        y = y.astype(np.float64)
        ny = np.linalg.norm(y)
        if ny == 0.0:
            continue
        raw.append(np.linalg.norm(pred - y) / ny)
        a = float(pred @ y) / max(float(pred @ pred), 1e-300)
        scaled.append(np.linalg.norm(a * pred - y) / ny)
    return float(np.mean(raw)), float(np.mean(scaled))


def train(sources, epochs: int, lr: float, batch: int = 8, seed: int = 1234, grid: int = 0, log=print, init=None):
    """Record the sources' pairs (restricted to grid^3 when grid > 0), then
    train with Adam under a cosine step-size decay, from random weights or
    from the weights at `init`; returns the network, the per-epoch history
    and the pair count."""
    from vapor_amd import mac3d

    m = grid or sources[0][0]
    capacity = sum(scenes * steps * (1 if fam == "simple" else 2) for _, _, scenes, steps, fam in sources)
    data = mac3d.Dataset(m, capacity)
    for n, first, scenes, steps, fam in sources:
        record_pairs(mac3d, data, n, range(first, first + scenes), steps, fam, log)
    pairs = len(data)
    held = []
    if grid:
        for vn, first, count, vsteps, vfam in VALIDATION:
            v = mac3d.Dataset(m, count * vsteps * (1 if vfam == "simple" else 2))
            record_pairs(mac3d, v, vn, range(first, first + count), vsteps, vfam)
            held.append((f"{vn}^3 {vfam}", v))
    net = mac3d.UNet(seed)
    if init:
        net.load(Path(init))
    net.set_grid(grid)
    history = []
    for e in range(epochs):
        t1 = time.perf_counter()
        # cosine decay of the step size over the epochs
        rate = lr * 0.5 * (1.0 + np.cos(np.pi * e / max(epochs, 1)))
        loss = net.train(data, 1, batch, float(rate), seed=e)
        entry = {"epoch": e, "lr": round(float(rate), 7), "train_loss": round(loss, 6),
                 "seconds": round(time.perf_counter() - t1, 1)}
        for name, v in held:
            raw, scaled = validation_error(net, v)
            entry[f"val_rel_err {name}"] = round(raw, 4)
            entry[f"val_rel_err_scaled {name}"] = round(scaled, 4)
        history.append(entry)
        log(json.dumps(entry))
    data.close()
    for _, v in held:
        v.close()
    return net, history, pairs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--protocol", choices=sorted(PROTOCOLS), default="coarse32")
    ap.add_argument("--epochs", type=int, default=0, help="epochs (default: the protocol's)")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--scale", type=float, default=1.0, help="multiplies every source's scene count")
    ap.add_argument("--out", type=str, default=None)
    ap.add_argument("--init", type=str, default=None,
                    help="continue from these weights (their record's history is kept in the new record)")
    args = ap.parse_args()
    sources, epochs, grid = PROTOCOLS[args.protocol]
    sources = [(n, first, max(1, int(round(scenes * args.scale))), steps, fam)
               for n, first, scenes, steps, fam in sources]
    epochs = args.epochs or epochs
    log = lambda *a: print(*a, flush=True, file=sys.stderr)  # noqa: E731
    t0 = time.perf_counter()
    net, history, pairs = train(sources, epochs, args.lr, args.batch, grid=grid, log=log, init=args.init)
    earlier = []
    if args.init and Path(args.init).with_suffix(".json").exists():
        prior = json.loads(Path(args.init).with_suffix(".json").read_text())
        earlier = [{"run": "earlier", "epochs": prior.get("epochs"), "lr": prior.get("lr"),
                    "history": prior.get("history")}] + list(prior.get("earlier", []))
    default = {"coarse32": coarse_weights(), "pairs64": warm_start_weights()}
    out = Path(args.out) if args.out else default.get(args.protocol, PACKAGE / "weights" / f"unet_{args.protocol}.bin")
    out.parent.mkdir(parents=True, exist_ok=True)
    net.save(out)
    record = {"weights": out.name, "protocol": args.protocol, "grid": grid,
              "sources": [{"n": n, "seeds": [first, first + scenes - 1], "steps": steps, "family": fam}
                          for n, first, scenes, steps, fam in sources],
              "pairs": pairs, "epochs": epochs, "lr": args.lr, "lr_schedule": "cosine", "batch": args.batch,
              "pre_sweeps": 32 if grid else 0, "history": history, "seconds": round(time.perf_counter() - t0, 1)}
    if args.init:
        record["init"] = Path(args.init).name
        record["earlier"] = earlier
    out.with_suffix(".json").write_text(json.dumps(record, indent=1) + "\n")
    print(json.dumps(record))


if __name__ == "__main__":
    main()
