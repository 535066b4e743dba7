"""Benchmark: PBD substeps/s of the drape256 workload on one MI355X.

Workload (one GPU): 256x256 cloth (65,536 vertices, 130,050 FEM triangles,
194,565 dihedral bends), side 2 m, dropped from 0.5 m onto a sphere of radius
0.3 given as a 64^3-cell grid SDF plus a ground plane; FEM (Young 1e4,
Poisson 0.3) + bending 0.01, 20 coloured Gauss-Seidel iterations, 4 substeps
per frame (dt 1/60), friction 0.3, thickness 2 mm, FP32, synthetic scene.

A step is one frame (scStepCloth: 4 substeps). W warm-up frames, then K
timed frames; the device is idle on both sides of the timed region (every
scStepCloth ends with a stream synchronisation). Prints one JSON line.

roofline: the timed frames (warm-up included) are replayed from a fresh scene
— the trajectory is bit-identical — with every kernel launch bracketed by HIP
events on the solver's stream; the dominant kernel's average duration comes
from those events, and achieved bandwidth is its algorithmic bytes per launch
(136 B per FEM triangle, 128 B per bend) over that duration. A rocprofv3
kernel trace of `bench.py --no-graph --no-cpu --copy-probe-mb 0` with the same
--steps/--warmup sees exactly these frames twice, so its per-kernel averages
are directly comparable. cpu_baseline: the CPU oracle (a single-threaded
restatement of the reference solver) on a bounded sample of the same workload.

roofline.traffic: HBM bytes per launch of the dominant kernel from the
committed rocprofv3 counter summary profiles/drape256_traffic.md (FETCH_SIZE
and WRITE_SIZE runs, each scaled by what the counter reads for a copy of known
size); used only when that summary was taken with the same launch phases per
sweep as this run, else null.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PBD substeps/sec at 256² cloth; wall time to 1e-4 residual, learned vs plain"
HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "drape256_traffic.md")
BYTES_PER_TRI = 136
BYTES_PER_BEND = 128
KERNEL_NAMES = {0: "k_fem", 1: "k_bend", 2: "k_contacts", 3: "k_detect", 4: "k_predict+k_velocity"}
WORKLOADS = {
    "drape256": "drape256: 256x256 cloth (65,536 verts, 130,050 FEM triangles, 194,565 bends) dropped onto a "
                "sphere (64^3 grid SDF) + ground; 20 coloured-GS iterations, 4 substeps/frame, friction 0.3, "
                "thickness 2 mm",
    "mesh512": "mesh512: 512x512 cloth (262,144 verts, 522,242 FEM triangles), height perturbed by three "
               "sinusoids (0.05 m), dropped onto an icosphere mesh (subdivision 6: 40,962 verts, 81,920 "
               "triangles, r 0.4) + ground; vertex, edge and triangle mesh contacts, continuous self "
               "collision (LBVH + FP64-cubic CCD); 10 iterations, 4 substeps/frame, friction 0.3, thickness 2 mm",
    "batch512": "batch512: 512 independent cloths (grid side U{16..128}, Young logU[1e3,1e5], Poisson "
                "U[0.1,0.45], bending logU[1e-3,1e-1], 1-4 pinned boundary vertices), each over its own sphere "
                "(r U[0.2,0.5], 32^3 grid SDF); one launch per substep for all cloths (one workgroup per cloth), "
                "graph-replayed, launch autotuned; 10 iterations, 4 substeps/frame, seed 4",
}
BYTES_PER_CONTACT = 92
BYTES_PER_VERTEX_PV = 64
BYTES_PER_SDF_DETECT = 204


def run_batch(args):
    """batch512: every step is one frame of all 512 cloths (4 batch launches)."""
    import numpy as np

    import solidcloth_amd as sc
    from solidcloth_amd import scenes as S

    scenes = S.batch_scenes(512)
    t0 = time.perf_counter()
    built = [S.build_patch(s, sc) for s in scenes]
    batch = sc.ClothBatch([p for _, p in built])
    batch.SetUseGraph(not args.no_graph)
    print(f"batch512: built {len(built)} patches in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    dt = scenes[0].dt
    warmup = max(args.warmup, 7)  # the autotuner times its 6 candidates on the first steps
    for _ in range(warmup):
        batch.Step(dt)
    dev_ms, launches = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.Step(dt)
        ms, n = batch.last_step_time()
        dev_ms += ms
        launches += n
    elapsed = time.perf_counter() - t0
    sub = scenes[0].substeps
    value = args.steps * sub / elapsed
    V, T, B = S.batch_counts(scenes)
    its = scenes[0].iterations
    passes = (its + 1) // 2
    contacts = sum(p.GetModel().GetNumContacts() for _, p in built)  # last substep's
    nonfinite = sum(int(not np.isfinite(p.GetModel().positions()).all()) for _, p in built)
    proj = its * T + passes * B + its * contacts
    bytes_launch = (BYTES_PER_VERTEX_PV + BYTES_PER_SDF_DETECT) * V + 136 * its * T + 128 * passes * B \
        + BYTES_PER_CONTACT * its * contacts
    avg_us = 1000.0 * dev_ms / max(1, launches)
    achieved = bytes_launch / (avg_us * 1e-6) / 1e9
    bs, stage, tuned = batch.GetLaunch()
    out = {
        "metric": "PBD substeps/sec, batch512 (every substep advances all 512 cloths)",
        "value": round(value, 2),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS["batch512"], "step": "one frame = 4 substeps of every cloth",
                   "graph_replay": not args.no_graph,
                   "cloths": len(scenes), "vertices": V, "triangles": T, "bends": B,
                   "launch": {"block_size": bs, "lds_staging": bool(stage), "autotuned": tuned},
                   "tuning_ms_per_step": [[b, s_, round(ms, 3)] for b, s_, ms in batch.tuning()]},
        "cloth_substeps_per_s": round(value * len(scenes), 1),
        "constraint_projections_per_s": round(proj * value, 1),
        "vertex_contacts_last_substep": contacts,
        "cloths_with_nonfinite_positions": nonfinite,
        "roofline": {
            "kernel": "k_batch_substep",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "avg_launch_us": round(avg_us, 3),
            "algorithmic_bytes_per_launch": int(bytes_launch),
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(scenes, args.cpu_budget)
    print(json.dumps(out))


def cpu_baseline_batch(scenes, budget_s):
    """The oracle steps the cloths one after another (one frame each) until the
    budget is spent; the batch rate is extrapolated by constraint count."""
    import oracle

    its = scenes[0].iterations
    passes = (its + 1) // 2

    def work(s):
        return its * 2 * (s.n - 1) ** 2 + passes * (3 * (s.n - 1) ** 2 - 2 * (s.n - 1))

    done, k, el = 0, 0, 0.0
    while k < len(scenes) and el < budget_s:
        o = oracle.from_scene(scenes[k])  # set-up is not timed
        t1 = time.perf_counter()
        o.step(scenes[k].dt)
        el += time.perf_counter() - t1
        done += work(scenes[k])
        k += 1
    total = sum(work(s) for s in scenes)
    per_substep_s = el / max(1, done) * total / scenes[0].substeps
    return {"value": round(1.0 / per_substep_s, 4), "unit": "substeps/s", "cores": 1, "kind": "port",
            "sample": f"{k} of {len(scenes)} cloths, one frame each, single-threaded oracle; the whole batch's "
                      "rate extrapolated by constraint-projection count"}


def run_gpu(scene, steps, warmup, use_graph=True, sweep=1):
    """The timed run: `warmup` frames, then `steps` frames under the clock."""
    import solidcloth_amd as sc
    from solidcloth_amd import scenes as S

    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetUseGraph(use_graph)
    m.SetSweep(sweep)
    for _ in range(warmup):
        patch.Step(scene.dt)
    t0 = time.perf_counter()
    for _ in range(steps):
        patch.Step(scene.dt)
    t1 = time.perf_counter()
    return t1 - t0, m.stats(), (world, patch, m)


def run_events(scene, frames, sweep=1):
    """The same frames again from a fresh scene (the trajectory is bit-identical),
    every kernel launch bracketed by HIP events on the solver's stream."""
    import solidcloth_amd as sc
    from solidcloth_amd import scenes as S

    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSweep(sweep)
    m.enable_kernel_timing(True)
    for _ in range(frames):
        patch.Step(scene.dt)
    timing = {k: m.kernel_timing(k) for k in KERNEL_NAMES}
    m.enable_kernel_timing(False)
    return timing


LEARNED_KS = (1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20)
NO_JUDGE = float("inf")


def learned_vs_plain(scene, start_frames=40, window=10, judged_frames=20, fixed_ks=(4, 6)):
    """The learned warm start against the plain solver on frames
    `start_frames`.. of the scene (trained parameters from
    solidcloth_amd/data/warmstart_mlp.npy; the scene is not among the
    training scenes).

    Per frame, from one saved state: the plain frame (the scene's iteration
    count) gives the target residual R and its device time. The warm-started
    frame runs with a fixed iteration count k (the production mode, no judge)
    and passes when its residual is at most R; the headline ratio is its
    device time over the plain time, failing frames included. The k-search
    (k = 1, 2, 3, ... until R is reached) is reported with every attempt
    charged, and the plain frame added when no k reaches R. Then the same
    frames are stepped with the per-step plain judge (each step runs plain
    and warm-started from the same start and keeps the better) and timed
    against plain frames: a verification mode, dearer than plain by design."""
    import numpy as np

    import solidcloth_amd as sc
    from solidcloth_amd import scenes as S
    from solidcloth_amd import warmstart_train as WT

    ws = WT.load()
    if ws is None:
        return {"status": "no trained parameters: run python -m solidcloth_amd.warmstart_train"}
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetWarmStart(ws, "track")
    for _ in range(start_frames):
        patch.Step(scene.dt)
    its = scene.iterations

    def frame(setup):
        m.restore_state()
        setup()
        patch.Step(scene.dt)
        return m.last_step_ms(), m.residual(), m.residual_terms()

    def plain(k):
        return lambda: (m.SetWarmStart(ws, "track"), m.SetNumIterations(k))

    def warm(k):
        return lambda: (m.SetWarmStart(ws, "solve", k), m.SetWarmStartFallback(NO_JUDGE))

    rows = []
    replay_ok = True
    for f in range(window):
        m.save_state()
        tp, R, tR = frame(plain(its))
        P20 = m.positions()
        _, r1, _ = frame(plain(1))
        _, r4, _ = frame(plain(4))
        m.SetNumIterations(its)
        row = {"t_plain": tp, "R": R, "shear_plain": tR[1], "r_plain1": r1, "r_plain4": r4, "fixed": {}}
        for k in fixed_ks:
            tw, rw, tW = frame(warm(k))
            row["fixed"][k] = (tw, rw, tW[1])
        charged, found = 0.0, None
        for k in LEARNED_KS:
            tk, rk, _ = frame(warm(k))
            charged += tk
            if rk <= R:
                found = k
                break
        if found is None:
            charged += tp
        row["k_search"] = (found, charged)
        rows.append(row)
        # continue along the plain trajectory
        m.restore_state()
        m.SetWarmStart(ws, "track")
        patch.Step(scene.dt)
        replay_ok = replay_ok and bool(np.array_equal(m.positions(), P20))
    t_plain = sum(r["t_plain"] for r in rows)
    fixed = {}
    for k in fixed_ks:
        passes = [bool(r["fixed"][k][1] <= r["R"]) for r in rows]
        fixed[str(k)] = {
            "frames_passing": int(sum(passes)),
            "per_frame_pass": passes,
            "warm_residual_mean": float(np.mean([r["fixed"][k][1] for r in rows])),
            "warm_shear_rms_mean": float(np.mean([r["fixed"][k][2] for r in rows])),
            "warm_ms_per_frame": float(np.mean([r["fixed"][k][0] for r in rows])),
            "learned_over_plain_time": float(sum(r["fixed"][k][0] for r in rows) / t_plain),
        }
    found = [r["k_search"][0] for r in rows if r["k_search"][0] is not None]
    # the per-step plain judge against plain frames over the same span
    m.save_state()
    m.SetWarmStart(ws, "track")
    plain_ms = 0.0
    for _ in range(judged_frames):
        patch.Step(scene.dt)
        plain_ms += m.last_step_ms()
    m.restore_state()
    m.SetWarmStart(ws, "solve", fixed_ks[0])
    m.SetWarmStartFallback(-1.0)
    judged_ms, kept = 0.0, 0
    for _ in range(judged_frames):
        patch.Step(scene.dt)
        judged_ms += m.last_step_ms()
        kept += 0 if m.warm_stats()["last_plain"] else 1
    return {
        "frames": f"{start_frames}..{start_frames + window - 1}",
        "residual": "RMS of warp and weft errors and bend errors, FP64 (shear reported apart)",
        "target": "per frame, the residual R of the plain solver's iteration count from the same state",
        "plain_over_plain1_residual": float(np.mean([r["R"] / r["r_plain1"] for r in rows])),
        "plain_iterations": its,
        "plain_residual_mean": float(np.mean([r["R"] for r in rows])),
        "plain_shear_rms_mean": float(np.mean([r["shear_plain"] for r in rows])),
        "plain1_residual_mean": float(np.mean([r["r_plain1"] for r in rows])),
        "plain4_residual_mean": float(np.mean([r["r_plain4"] for r in rows])),
        "plain_ms_per_frame": t_plain / len(rows),
        "fixed_k": fixed,
        "headline": {"iterations": fixed_ks[0], "learned_over_plain_time": fixed[str(fixed_ks[0])]["learned_over_plain_time"],
                     "frames_passing": fixed[str(fixed_ks[0])]["frames_passing"], "frames": len(rows)},
        "k_search_charged": {"frames_reaching": len(found), "iterations_mean": float(np.mean(found)) if found else None,
                             "learned_over_plain_time": float(sum(r["k_search"][1] for r in rows) / t_plain)},
        "plain_replay_identical": replay_ok,
        "judged_mode": {"frames": judged_frames, "iterations": fixed_ks[0], "judge": "per-step plain step",
                        "plain_ms": round(plain_ms, 3), "judged_ms": round(judged_ms, 3),
                        "ratio": round(judged_ms / plain_ms, 4), "warm_steps_kept": kept,
                        "fallbacks": m.warm_stats()["fallbacks"]},
    }


def measured_traffic(kernel, phases):
    """HBM bytes per launch of `kernel` ("k_fem" / "k_bend") from the counter
    summary's JSON block, or None when absent or taken with other phases."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        d = json.loads(block)
    except (OSError, IndexError, ValueError):
        return None
    if d.get("launch_phases_per_pass") != phases:
        return None
    k = d.get("kernels", {}).get(kernel)
    return k.get("hbm_bytes_per_launch") if k else None


def render_sync(patch, with_cpu, calls=20):
    """The render-mesh sync (scUpdateClothMesh: vertices, angle-weighted
    normals, face normals, copied to the host mesh buffers) per call on the
    device, and the oracle's host restatement of the same normals
    (Mesh::ComputeNormals on the same vertices) beside it."""
    patch.UpdateMesh()  # builds the vertex-triangle lists once
    t0 = time.perf_counter()
    for _ in range(calls):
        patch.UpdateMesh()
    gpu_ms = 1000.0 * (time.perf_counter() - t0) / calls
    verts, _, idx = patch.mesh_arrays()
    out = {"vertices": int(len(verts)), "triangles": int(len(idx)), "gpu_ms_per_call": round(gpu_ms, 3),
           "includes": "kernels + device-to-host copies of vertices and normals (1.6 MB)"}
    if with_cpu:
        import oracle

        oracle.mesh_normals(verts, idx)
        t0 = time.perf_counter()
        n = 3
        for _ in range(n):
            oracle.mesh_normals(verts, idx)
        out["cpu_oracle_ms_per_call"] = round(1000.0 * (time.perf_counter() - t0) / n, 3)
        out["cpu_cores"] = 1
    return out


def cpu_baseline(scene, budget_s):
    import oracle

    o = oracle.from_scene(scene)
    # one frame to estimate, then as many frames as fit the budget
    t0 = time.perf_counter()
    o.step(scene.dt)
    one = time.perf_counter() - t0
    frames = max(1, min(50, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(frames):
        o.step(scene.dt)
    el = time.perf_counter() - t0
    return frames * scene.substeps / el, frames


def cpu_baseline_from_state(scene, pos, vel, budget_s):
    """Substeps/s of the oracle with its grid broad phases (the same contacts
    as its exhaustive scans, proven by the CPU tests), started from the state
    the GPU run ended in, so that the sample has the timed region's contacts."""
    import oracle

    o = oracle.from_scene(scene)
    o.set(accel=1, numSteps=1)
    o.put("pos", pos)
    o.put("prev", pos)
    o.put("vel", vel)
    h = scene.dt / scene.substeps
    t0 = time.perf_counter()
    o.step(h)
    one = time.perf_counter() - t0
    n = max(1, min(20, int(budget_s / max(one, 1e-6)) - 1))
    t0 = time.perf_counter()
    for _ in range(n):
        o.step(h)
    el = time.perf_counter() - t0
    c = o.counts()["C"] + len(o.self_contacts("vt")[0]) + len(o.self_contacts("ee")[0])
    return n / el, n, c


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels directly instead of replaying the captured substep graph")
    ap.add_argument("--sweep", type=int, default=3, choices=[0, 1, 3],
                    help="0: one launch per colour; 1: slab phases staged in LDS; 3: triangles on tiles with "
                         "halos (one launch per FEM pass), bends on slabs")
    ap.add_argument("--copy-probe-mb", type=int, default=1024,
                    help="size of the streaming-copy bandwidth probe (0 disables)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--workload", default="drape256", choices=["drape256", "mesh512", "batch512"],
                    help="drape256: the metric's workload; mesh512: the triangle-mesh obstacle path "
                         "(reported beside it, not the metric)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is design-only "
                                   "(see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    from solidcloth_amd import scenes as S

    if args.workload == "batch512":
        run_batch(args)
        return
    scene = S.DRAPE256 if args.workload == "drape256" else S.MESH512
    elapsed, stats, keep = run_gpu(scene, args.steps, args.warmup, use_graph=not args.no_graph, sweep=args.sweep)
    m = keep[2]
    event_frames = args.warmup + args.steps
    timing = run_events(scene, event_frames, sweep=args.sweep)
    substeps = args.steps * scene.substeps
    value = substeps / elapsed
    T = scene.num_triangles
    B = int(m.GetNumBends())
    its = scene.iterations
    bend_passes = (its + 1) // 2
    proj_per_substep = its * T + bend_passes * B

    # dominant kernel by total event time
    dom = max(timing, key=lambda k: timing[k][1])
    launches, total_ms = timing[dom]
    avg_us = 1000.0 * total_ms / max(1, launches)
    frames_t = event_frames
    if dom == 0:
        units = T * its * scene.substeps * frames_t
        bytes_per_launch = BYTES_PER_TRI * units / launches
    elif dom == 1:
        units = B * bend_passes * scene.substeps * frames_t
        bytes_per_launch = BYTES_PER_BEND * units / launches
    else:
        bytes_per_launch = None
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9 if bytes_per_launch else None

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload],
            "step": "one frame = 4 substeps",
            "iterations": its,
            "substeps_per_frame": scene.substeps,
            "graph_replay": not args.no_graph,
            "sweep": {0: "colour launches", 1: "slab phases (LDS)", 3: "FEM tiles with halos, bend slab phases (LDS)"}[args.sweep],
            "launch_phases_per_pass": {"fem": len(m.colour_offsets("tri")) - 1,
                                       "bend": len(m.colour_offsets("bend")) - 1},
        },
        "constraint_projections_per_s": round(proj_per_substep * value, 1),
        "solver_iterations_last_substep": stats["iterations"],
        "kernel_avg_us": {KERNEL_NAMES[k]: round(1000.0 * timing[k][1] / max(1, timing[k][0]), 3)
                          for k in KERNEL_NAMES},
        "roofline": {
            "kernel": KERNEL_NAMES[dom],
            "bound": "hbm",
            "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": None,
            "avg_launch_us": round(avg_us, 3),
            "algorithmic_bytes_per_launch": round(bytes_per_launch) if bytes_per_launch else None,
        },
    }
    if args.workload == "drape256" and dom in (0, 1):
        t = measured_traffic(KERNEL_NAMES[dom], out["config"]["launch_phases_per_pass"])
        out["roofline"]["traffic"] = round(t) if t else None
    if args.copy_probe_mb > 0:
        import solidcloth_amd as sc

        nbytes = args.copy_probe_mb << 20
        ms = sc.lib().scBandwidthProbe(nbytes, 20)
        if ms > 0:
            out["copy_probe"] = {"bytes_copied": nbytes, "avg_ms": round(ms, 4),
                                 "GB/s (read+write)": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}
    if args.workload == "drape256":
        out["render_mesh_sync"] = render_sync(keep[1], not args.no_cpu)
    if args.workload == "drape256" and not args.no_learned:
        out["learned_vs_plain"] = learned_vs_plain(scene)
    if args.workload != "drape256":
        out["metric"] = "PBD substeps/sec, " + args.workload
        out["cpu_baseline"] = None
        if not args.no_cpu:
            pos = m.positions().astype("float64")
            vel = m.velocities().astype("float64")
            cpu_val, n, contacts = cpu_baseline_from_state(scene, pos, vel, args.cpu_budget)
            out["cpu_baseline"] = {
                "value": round(cpu_val, 4),
                "unit": "substeps/s",
                "cores": 1,
                "kind": "port",
                "sample": f"{n} substeps of mesh512 from the state the GPU run ended in ({contacts} contacts "
                          "in the last), single-threaded oracle restatement with grid broad phases",
            }
    elif not args.no_cpu:
        cpu_val, frames = cpu_baseline(scene, args.cpu_budget)
        out["cpu_baseline"] = {
            "value": round(cpu_val, 3),
            "unit": "substeps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{frames} frames ({frames * scene.substeps} substeps) of drape256 from the initial state, "
                      "single-threaded oracle restatement",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
