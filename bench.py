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
are directly comparable. cpu_baseline: the CPU oracle (the restatement of the
reference solver: serial Gauss-Seidel, OpenMP detection) on a bounded sample
of the same workload, at the host's thread count and at one thread, with the
CPU model.

mesh512: the self-colliding 512x512 cloth over all its 300 frames on the GPU,
with two CPU baselines: the algorithm-matched path (the oracle's grid broad
phase + the reference's CCD) and the reference's live brute-force self path
(sampled queries of substep 1, scaled), and the GPU/CPU ratios against 50x.

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
    balance = batch_load_balance(batch, dt)
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
        "load_balance": balance,
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
        out["cpu_baseline"] = cpu_baseline_batch(scenes, args.cpu_budget, host_cpu())
    print(json.dumps(out))


def batch_load_balance(batch, dt):
    """One more (untimed) step with per-workgroup clocks: how the last
    substep launch's cloths filled the GPU. busy_fraction = the workgroups'
    summed run time over (launch span x the most workgroups seen running at
    once); the launch order is largest cloth first."""
    import numpy as np

    batch.RecordWorkgroupTimes(True)
    batch.Step(dt)
    wt = batch.workgroup_times()
    batch.RecordWorkgroupTimes(False)
    st, en = wt["start_us"], wt["end_us"]
    dur = en - st
    span = float(en.max() - st.min())
    ev = sorted([(t, 1) for t in st] + [(t, -1) for t in en], key=lambda e: (e[0], e[1]))
    running = concurrent = 0
    for _, d in ev:
        running += d
        concurrent = max(concurrent, running)
    units = len({tuple(u) for u in wt["unit"].tolist()})
    return {
        "cloths": int(len(st)),
        "units_used": units,
        "max_concurrent_workgroups": concurrent,
        "span_us": round(span, 2),
        "busy_fraction": round(float(dur.sum()) / (span * concurrent), 4),
        "cloth_us": {"min": round(float(dur.min()), 2), "mean": round(float(dur.mean()), 2),
                     "max": round(float(dur.max()), 2)},
        "last_start_us": round(float(st.max() - st.min()), 2),
        "note": "last substep launch of one extra step, per-workgroup s_memrealtime clocks",
    }


def cpu_baseline_batch(scenes, budget_s, cpu):
    """The oracle steps the cloths (one frame each) until the budget is spent,
    one after another on one thread, then on the host's thread count with the
    cloths spread over a thread pool (one cloth per worker at a time, one
    OpenMP thread each; set-up is not timed); the batch rate is extrapolated
    by constraint count."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle

    its = scenes[0].iterations
    passes = (its + 1) // 2

    def work(s):
        return its * 2 * (s.n - 1) ** 2 + passes * (3 * (s.n - 1) ** 2 - 2 * (s.n - 1))

    total = sum(work(s) for s in scenes)
    sub = scenes[0].substeps
    done, k, el = 0, 0, 0.0
    oracle.set_threads(1)
    while k < len(scenes) and el < budget_s / 2:
        o = oracle.from_scene(scenes[k])  # set-up is not timed
        t1 = time.perf_counter()
        o.step(scenes[k].dt)
        el += time.perf_counter() - t1
        done += work(scenes[k])
        k += 1
    single = 1.0 / (el / max(1, done) * total / sub)
    # the same cloths, then more, on the thread pool: as many as one thread
    # stepped, times the thread count (bounded by the batch)
    n = min(len(scenes), max(k, k * cpu["threads"]))
    cloths = [oracle.from_scene(s) for s in scenes[:n]]

    def step(i):
        oracle.set_threads(1)
        cloths[i].step(scenes[i].dt)

    with ThreadPoolExecutor(max_workers=cpu["threads"]) as pool:
        t1 = time.perf_counter()
        list(pool.map(step, range(n)))
        wall = time.perf_counter() - t1
    parallel = 1.0 / (wall / sum(work(s) for s in scenes[:n]) * total / sub)
    return {"value": round(parallel, 4), "unit": "substeps/s", "cores": cpu["threads"], "kind": "port",
            "sample": f"{n} of {len(scenes)} cloths, one frame each, oracle restatement, cloths spread over "
                      f"{cpu['threads']} threads; the whole batch's rate extrapolated by constraint-projection count",
            "single_thread": {"value": round(single, 4), "sample": f"{k} cloths one after another"},
            "cpu_model": cpu["model"], "logical_cpus": cpu["logical_cpus"], "available_cpus": cpu["available_cpus"]}


def run_gpu(scene, steps, warmup, use_graph=True, sweep=1, slab_size=0):
    """The timed run: `warmup` frames, then `steps` frames under the clock."""
    import solidcloth_amd as sc
    from solidcloth_amd import scenes as S

    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetUseGraph(use_graph)
    m.SetSweep(sweep, slab_size)
    for _ in range(warmup):
        patch.Step(scene.dt)
    t0 = time.perf_counter()
    for _ in range(steps):
        patch.Step(scene.dt)
    t1 = time.perf_counter()
    return t1 - t0, m.stats(), (world, patch, m)


def run_events(scene, frames, sweep=1, slab_size=0):
    """The same frames again from a fresh scene (the trajectory is bit-identical),
    every kernel launch bracketed by HIP events on the solver's stream."""
    import solidcloth_amd as sc
    from solidcloth_amd import scenes as S

    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSweep(sweep, slab_size)
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


def measured_traffic(kernel, phases, sweep):
    """HBM bytes per launch of `kernel` ("k_fem" / "k_bend") from the counter
    summary's JSON block, or None when absent or taken with other phases or
    another sweep mode."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        d = json.loads(block)
    except (OSError, IndexError, ValueError):
        return None
    if d.get("launch_phases_per_pass") != phases or d.get("sweep_mode") != sweep:
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


def host_cpu():
    """The host's CPU model and the thread count the CPU baselines use: the
    OpenMP thread count the environment sets (the host's share of cores), else
    every core this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else avail
    return {"model": model, "logical_cpus": os.cpu_count(), "available_cpus": avail,
            "threads": max(1, min(threads, avail))}


def _oracle_substeps(o, h, budget_s, limit=50):
    """Times substeps of an oracle cloth until the budget is spent (at least one)."""
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(h)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= limit:
            return n / el, n


def cpu_baseline(scene, budget_s, cpu):
    """drape256 on the oracle from the initial state, at the host's thread
    count and at one thread (Gauss-Seidel is serial, as the reference's; the
    detection loops use OpenMP)."""
    import oracle

    h = scene.dt / scene.substeps
    out = {}
    for threads in (cpu["threads"], 1):
        oracle.set_threads(threads)
        o = oracle.from_scene(scene)
        o.set(numSteps=1)
        out[threads] = _oracle_substeps(o, h, budget_s / 2)
    oracle.set_threads(1)
    (vn, nn), (v1, n1) = out[cpu["threads"]], out[1]
    return {"value": round(vn, 3), "unit": "substeps/s", "cores": cpu["threads"], "kind": "port",
            "sample": f"{nn} substeps of drape256 from the initial state, oracle restatement (serial Gauss-Seidel, "
                      f"OpenMP detection) at {cpu['threads']} threads",
            "single_thread": {"value": round(v1, 3), "sample": f"{n1} substeps from the initial state"},
            "cpu_model": cpu["model"], "logical_cpus": cpu["logical_cpus"], "available_cpus": cpu["available_cpus"]}


def ccd_path_baseline(scene, pos, vel, budget_s, cpu):
    """mesh512 on the oracle with its grid broad phases (the same contacts as
    its exhaustive scans) and the reference's CCD, from the state the GPU run
    ended in, so that the sample has the timed region's contacts; at the
    host's thread count, then at one thread continuing from there."""
    import oracle

    o = oracle.from_scene(scene)
    o.set(accel=1, numSteps=1)
    o.put("pos", pos)
    o.put("prev", pos)
    o.put("vel", vel)
    h = scene.dt / scene.substeps
    res = {}
    for threads in (cpu["threads"], 1):
        oracle.set_threads(threads)
        rate, n = _oracle_substeps(o, h, budget_s / 2, limit=20)
        c = o.counts()["C"] + len(o.self_contacts("vt")[0]) + len(o.self_contacts("ee")[0])
        res[threads] = (rate, n, c)
    oracle.set_threads(1)
    (vn, nn, cn), (v1, n1, c1) = res[cpu["threads"]], res[1]
    return {"value": round(vn, 4), "unit": "substeps/s", "cores": cpu["threads"], "kind": "port",
            "sample": f"{nn} substeps of mesh512 from the state the GPU run ended in ({cn} contacts in the last), "
                      f"oracle restatement with grid broad phases and the reference's CCD at {cpu['threads']} "
                      "threads",
            "single_thread": {"value": round(v1, 4), "sample": f"{n1} further substeps ({c1} contacts in the last)"}}


LIVE_SAMPLE = (64, 128)  # vertex and edge queries timed per thread count


def live_path_baseline(pos0, tris, rad_tol, cpu):
    """The reference's live self path at substep 1 (brute force over every
    triangle per vertex and every later edge per edge; ClosestPointsBetweenMeshes
    .cpp:538-592), timed on evenly spaced vertex and edge queries and scaled to
    all of them: per-vertex cost is the same for every vertex, and an edge's
    cost falls linearly with its index, so midpoint-spaced edges sum to the
    whole. The compiled reference's own functions in the reference's OpenMP
    loop structure (oracle/_ref) when present, else the restatement."""
    import ctypes as C

    import numpy as np

    import oracle

    v = np.ascontiguousarray(pos0, dtype=np.float32)
    idx = np.ascontiguousarray(tris, dtype=np.uint32).ravel()
    nv, nt = len(v), len(idx) // 3
    ne = len(oracle.construct_edges(idx))
    kv, ke = LIVE_SAMPLE
    vq = ((np.arange(kv) + 0.5) * nv / kv).astype(np.int32)
    eq = ((np.arange(ke) + 0.5) * ne / ke).astype(np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref_geometry.so")
    runs = {}
    kind = "reference" if os.path.exists(ref_path) else "port"
    if kind == "reference":
        R = C.CDLL(ref_path)
        for threads in (cpu["threads"], 1):
            R.ref_set_threads(C.c_int(threads))
            cnt = np.zeros(2, np.int32)
            sec = np.zeros(2, np.float64)
            R.ref_live_self_sample(p(v), C.c_int(nv), p(idx), C.c_int(idx.size), C.c_float(rad_tol), p(vq),
                                   C.c_int(kv), p(eq), C.c_int(ke), p(cnt), p(sec))
            runs[threads] = (sec[0], sec[1])
    live = oracle.LiveSelf(v, idx)
    port = {}
    for threads in ((cpu["threads"], 1) if kind == "port" else (1,)):
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        live.vertex_queries(rad_tol, vq)
        t1 = time.perf_counter()
        live.edge_queries(eq)
        port[threads] = (t1 - t0, time.perf_counter() - t1)
    oracle.set_threads(1)
    if kind == "port":
        runs = port

    def per_substep(t):
        return t[0] * nv / kv + t[1] * ne / ke

    sn, s1 = per_substep(runs[cpu["threads"]]), per_substep(runs[1])
    out = {"value": 1.0 / sn, "unit": "substeps/s", "cores": cpu["threads"], "kind": kind,
           "seconds_per_substep": round(sn, 1),
           "sample": f"substep 1 of mesh512 (radius {rad_tol:.4f}): {kv} of {nv} vertex queries and {ke} of {ne} "
                     f"edge queries, scaled to all; {'compiled reference functions' if kind == 'reference' else 'restatement'}"
                     f" at {cpu['threads']} threads",
           "single_thread": {"seconds_per_substep": round(s1, 1)}}
    if kind == "reference":
        out["restatement_single_thread_seconds_per_substep"] = round(per_substep(port[1]), 1)
    return out


def mesh512_record(args, cpu):
    """The self-colliding 512^2 cloth: all configured frames on the GPU, then
    the CPU baselines of the algorithm-matched CCD path and of the reference's
    live self path, and the GPU / CPU ratios against the 50x target."""
    import solidcloth_amd as sc
    from solidcloth_amd import scenes as S

    scene = S.MESH512
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetUseGraph(not args.no_graph)
    m.SetSweep(args.sweep, args.slab_size)
    pos0 = m.positions().copy()
    t0 = time.perf_counter()
    patch.Step(scene.dt)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    most = {"vertices_vt": 0, "vertices_ee": 0}
    for f in range(1, scene.frames):
        patch.Step(scene.dt)
        if f % 50 == 0:
            print(f"mesh512: frame {f}/{scene.frames}, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if f % 10 == 0:
            sf = m.self_stats()
            for k in most:
                most[k] = max(most[k], sf[k])
    el = time.perf_counter() - t0
    value = (scene.frames - 1) * scene.substeps / el
    st = m.self_stats()
    pos = m.positions()
    out = {"workload": WORKLOADS["mesh512"], "frames": scene.frames, "value": round(value, 3), "unit": "substeps/s",
           "timed": f"frames 2-{scene.frames} (frame 1, with graph capture, took {1000 * first:.0f} ms)",
           "ms_per_frame": round(1000.0 * el / (scene.frames - 1), 3),
           "all_positions_finite": bool(__import__("numpy").isfinite(pos).all()),
           "self_collision": dict(st, most_contact_vertices_sampled=most)}
    if args.no_cpu:
        return out
    tris = m.triangles()
    ccd = ccd_path_baseline(scene, pos.astype("float64"), m.velocities().astype("float64"), args.cpu_budget, cpu)
    # the live path at substep 1: the mesh is the initial positions, radius
    # thickness + max(tolerance, 1.5 x the largest first prediction step)
    h = scene.dt / scene.substeps
    rad = scene.thickness + max(scene.tolerance, 1.5 * 9.8 * h * h)
    live = live_path_baseline(pos0, tris, rad, cpu)
    out["cpu_baseline"] = {"ccd_path": ccd, "live_path": live, "cpu_model": cpu["model"],
                           "logical_cpus": cpu["logical_cpus"], "available_cpus": cpu["available_cpus"]}
    rc, rl = value / ccd["value"], value / live["value"]
    out["gpu_over_cpu"] = {"ccd_path": round(rc, 1), "live_path": round(rl), "target": 50,
                           "met": bool(rc >= 50 and rl >= 50),
                           "ccd_path_single_thread": round(value / ccd["single_thread"]["value"], 1),
                           "live_path_single_thread": round(value * live["single_thread"]["seconds_per_substep"])}
    live["value"] = round(live["value"], 8)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels directly instead of replaying the captured substep graph")
    ap.add_argument("--sweep", type=int, default=4, choices=[0, 1, 3, 4],
                    help="0: one launch per colour; 1: slab phases staged in LDS; 3: triangles on tiles with "
                         "halos (one launch per FEM pass), bends on slabs; 4: the schedules of 3 with each "
                         "launch's data requested before its first colour")
    ap.add_argument("--slab-size", type=int, default=0,
                    help="slab size of the sweep (modes 3 and 4: the interiors of the bend slabs); 0: the "
                         "library's choice")
    ap.add_argument("--copy-probe-mb", type=int, default=1024,
                    help="size of the streaming-copy bandwidth probe (0 disables)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--workload", default="drape256", choices=["drape256", "mesh512", "batch512"],
                    help="drape256: the metric's workload; mesh512: the self-colliding triangle-mesh obstacle "
                         "path alone (its record is also part of the drape256 line)")
    ap.add_argument("--no-mesh512", action="store_true",
                    help="leave the self-colliding 512^2 record (300 frames + CPU baselines) out of the line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is design-only "
                                   "(see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    from solidcloth_amd import scenes as S

    if args.workload == "batch512":
        run_batch(args)
        return
    cpu = host_cpu()
    if args.workload == "mesh512":
        rec = mesh512_record(args, cpu)
        print(json.dumps({"metric": "PBD substeps/sec, mesh512 (self-colliding 512x512 cloth, all frames)",
                          "value": rec["value"], "unit": "substeps/s", "n_gpus": 1, "steps": S.MESH512.frames,
                          "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "f32", "data": "synthetic", "config": {"workload": rec["workload"]},
                          "mesh512": rec}))
        return
    scene = S.DRAPE256
    elapsed, stats, keep = run_gpu(scene, args.steps, args.warmup, use_graph=not args.no_graph, sweep=args.sweep,
                                     slab_size=args.slab_size)
    m = keep[2]
    event_frames = args.warmup + args.steps
    timing = run_events(scene, event_frames, sweep=args.sweep, slab_size=args.slab_size)
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
            "sweep": {0: "colour launches", 1: "slab phases (LDS)", 3: "FEM tiles with halos, bend slab phases (LDS)",
                      4: "FEM tiles with halos, bend slab phases (LDS), packed: all of a launch's data requested "
                         "before its first colour"}[args.sweep],
            "sweep_mode": args.sweep,
            "slab_size": args.slab_size,
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
        t = measured_traffic(KERNEL_NAMES[dom], out["config"]["launch_phases_per_pass"], args.sweep)
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
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget, cpu)
    if not args.no_mesh512:
        out["mesh512"] = mesh512_record(args, cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
