"""Training of the learned warm-start on trajectories of the plain solver.

Every substep of a training scene runs the plain solve with the network in
training mode: the features are taken at the inertial positions, and after
the sweeps the network takes one Adam step (BF16-MFMA forward and backward
kernels on the GPU) towards the correction the sweeps applied. The scenes
come from the three plain workloads' families (drape over a sphere SDF,
triangle-mesh obstacle with self collision, batched small cloths) with their
parameters varied, so that the drape256 scene itself is held out.

    python -m solidcloth_amd.warmstart_train --out solidcloth_amd/data/warmstart_mlp.npy

writes the trained parameters (5187 float32, the order of scWarmStartGetParams).
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "warmstart_mlp.npy")


def training_plan():
    """(scene, frames) pairs: about 5,000 substeps in all."""
    from . import scenes as S

    plan = []
    # drape family (the drape256 scene is held out): other sizes, spheres,
    # heights and stiffnesses
    for n, r, y, young, mu in ((128, 0.25, 0.45, 5e3, 0.3), (192, 0.35, 0.6, 2e4, 0.2),
                               (160, 0.3, 0.5, 1e4, 0.4), (256, 0.28, 0.55, 8e3, 0.3)):
        sc = S.DRAPE256.with_(n=n, sphere_radius=r, sdf_half=r + 0.1, offset=(0.0, y, 0.0), young=young,
                              friction=mu)
        plan.append((sc, 200))
    # triangle-mesh obstacle with self collision, at 256^2
    m = S.MESH512.with_(n=256, ccd=S.ccd_for_spacing(2.0 / 255), perturb_seed=7)
    plan.append((m, 60))
    # batched-workload cloths, stepped one by one
    for sc in S.batch_scenes(16, seed=21, n_range=(24, 128)):
        plan.append((sc, 25))
    return plan


def train(plan=None, seed=7, lr=1e-3, lr_final=3e-4, log=None, use_graph=True, epochs=1):
    """Trains a fresh network over the plan, `epochs` times (the plain
    trajectories are deterministic, so every epoch sees the same substeps);
    returns (network, substeps per epoch, seconds)."""
    import solidcloth_amd as sc
    from . import scenes as S

    plan = training_plan() if plan is None else plan
    ws = sc.WarmStart(seed)
    total = epochs * sum(f * s.substeps for s, f in plan)
    done = 0
    t0 = time.perf_counter()
    for k, (scene, frames) in enumerate(list(plan) * epochs):
        world, patch = S.build_patch(scene, sc)
        m = patch.GetModel()
        m.SetUseGraph(use_graph)
        m.SetWarmStart(ws, "train")
        for f in range(frames):
            # linear decay of the step size over the whole plan
            ws.SetLearningRate(lr + (lr_final - lr) * done / max(1, total))
            try:
                patch.Step(scene.dt)
            except sc.SolidClothError as e:  # e.g. a self-collision overflow: the scene ends there
                if log:
                    print(f"warm-start training: scene {scene.name} stopped at frame {f}: {e}", file=log, flush=True)
                break
            done += scene.substeps
        if log:
            print(f"warm-start training: scene {k + 1}/{len(plan) * epochs} ({scene.name}, n={scene.n}) done, "
                  f"{done} substeps, {time.perf_counter() - t0:.1f} s", file=log, flush=True)
        del patch, world
    return ws, done // epochs, time.perf_counter() - t0


def load(path=DEFAULT_WEIGHTS, seed=0):
    """A network with the saved parameters, or None when the file is absent."""
    import solidcloth_amd as sc

    if not os.path.exists(path):
        return None
    p = np.load(path, allow_pickle=False)
    ws = sc.WarmStart(seed)
    ws.set_params(p)
    return ws


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--epochs", type=int, default=3, help="passes over the same trajectories")
    ap.add_argument("--frames-scale", type=float, default=1.0,
                    help="fraction of each scene's frames (shorter runs, e.g. under a profiler)")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels directly instead of graph replay")
    args = ap.parse_args()
    plan = [(s, max(1, int(round(f * args.frames_scale)))) for s, f in training_plan()]
    ws, n, secs = train(plan=plan, seed=args.seed, log=sys.stdout, use_graph=not args.no_graph,
                          epochs=args.epochs)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, ws.params())
    print(f"trained on {n} substeps x {args.epochs} epochs in {secs:.1f} s; {ws.steps()} Adam steps; saved {args.out}", flush=True)


if __name__ == "__main__":
    main()
