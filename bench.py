"""Benchmark of the WCSPH dam-break step on one MI355X.

Workload (the project's WCSPH dam-break configuration): a 0.4 x 0.4 x 0.2 m
water column at dx = 5 mm (80 x 80 x 40 = 256,000 fluid particles) in a
1.6 x 0.4 x 0.6 m open tank with a 3-layer dummy shell, Wendland C2 with
h = 1.3 dx, rho0 = 1000, Tait EOS gamma = 7, c0 = 10 sqrt(2 g H), Riemann
pressure with limiter, dt = 0.2 h / c0.  One step is one iteration of the
solver loop: hash, sort, reorder, cell ranges, neighbour lists, then the
three pair passes (momentum + kick/drift, continuity + EOS, momentum + kick).
All state is resident in HBM before the timed region starts.

Prints one JSON line:
  value      = fluid particle-steps per second over the timed steps
  roofline   = the dominant kernel's algorithmic bytes / its mean launch time
               (HIP events on the scene stream, inside the timed region);
               traffic = its rocprofv3-measured HBM bytes per launch
  cpu_baseline = the FP64 oracle on the host cores over a bounded sample
  dem_collapse = the 4,096-ball DEM granular collapse, ball-steps per second,
               with its own oracle baseline
  serving_mix  = the batched-serving job stream, plain vs batched (and
               batched with autotuning on admission): throughput, p50/p99
  autotune     = the autotuner's trials on the dam break against the rules
  state_io     = the dam break's state exported to host memory and imported
               back (the cost of evicting and resuming a served job)
  cantilever_clumps, ice_cylinder_coupled = the bonded-beam-with-clumps and
               coupled ice-cylinder workloads
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s + HBM GB/s; batched-scene p50/p99 step latency vs plain"
HBM_PEAK_GBPS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def algorithmic_bytes(kernel: str, nf: int, pairs_ff: int, pairs_fd: int) -> float:
    """Compulsory HBM bytes of one launch: each array the launch needs is read
    or written once; neighbour re-reads served by caches are not counted.
    Layout: FluidPos {x,y,z f64; rho,p f32} 32 B, V {v f32 x3; phase} 16 B,
    FP64 velocity 24 B, drho 4 B, list entries 4 B, per-slot counts 8 B."""
    lists = 4 * (pairs_ff + pairs_fd) + 8 * nf
    if kernel == "sph_momentum_first":
        return nf * (80 + 72) + lists      # read P, V, v64, counts; write P', V', v64'
    if kernel == "sph_momentum_second":
        return nf * (80 + 40) + lists      # read P, V, v64, counts; write V', v64'
    if kernel == "sph_continuity":
        return nf * (60 + 36) + lists      # read P, V, drho, counts; write P', drho'
    if kernel == "sph_lists":
        return nf * 40 + lists             # read own x and its FP32 copy; write the lists and counts
    if kernel == "sph_reorder":
        return nf * 184                    # gather P, V, v64, drho, ids; write them and the FP32 copy
    return 0.0


PROFILE = ROOT / "profiles" / "wcsph_dam_break_256k_kernels.md"
# profile table rows are named by kernel; the bench names the launches
PROFILE_KERNEL = {"sph_lists": "k_build_lists", "sph_momentum_first": "k_momentum<true",
                  "sph_momentum_second": "k_momentum<false", "sph_continuity": "k_continuity",
                  "sph_reorder": "k_reorder"}


def profiled_traffic(name: str):
    """HBM bytes per launch of a step kernel, read from the committed rocprofv3
    summary (its FETCH_SIZE / WRITE_SIZE columns, already scaled by the copy
    calibration): (bytes, source) or (None, reason)."""
    key = PROFILE_KERNEL.get(name)
    if key is None or not PROFILE.exists():
        return None, "no profile row"
    for line in PROFILE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0].startswith(key):
            try:
                return (float(cells[4]) + float(cells[5])) * 1e6, f"{PROFILE.name}: {cells[0]}"
            except ValueError:
                return None, "unreadable profile row"
    return None, "no profile row"


# launches per WCSPH step of each kernel in the profile's table (2 radix passes)
STEP_KERNELS = {"k_hash": 1, "k_radix_hist": 2, "k_radix_tile_scan": 2, "k_radix_scatter": 2, "k_reorder": 1,
                "k_build_lists": 1, "k_momentum<true": 1, "k_continuity": 1, "k_momentum<false": 1}


def profiled_step_traffic():
    """HBM bytes of one dam-break step from the committed rocprofv3 summary:
    per-launch FETCH_SIZE + WRITE_SIZE bytes (copy-calibrated) of every step
    kernel times its launches per step, or None without the profile."""
    if not PROFILE.exists():
        return None
    total, seen = 0.0, set()
    for line in PROFILE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) != 6:
            continue
        for key, per_step in STEP_KERNELS.items():
            if cells[0].startswith(key) and key not in seen:
                try:
                    total += (float(cells[4]) + float(cells[5])) * 1e6 * per_step
                except ValueError:
                    return None
                seen.add(key)
    return total if seen == set(STEP_KERNELS) else None


def profiled_limiter(name: str, avg_us: float):
    """What bounds a step kernel, from the committed profile's counter table
    (SQ_WAVES, SQ_INSTS_VALU, SQ_WAIT_ANY / SQ_WAVE_CYCLES per dispatch) and
    this run's launch time: the SIMD time that VALU issue needs (a wave64
    instruction holds its SIMD 4 cycles; 1,024 SIMDs at 2.4 GHz) against the
    launch, and the share of wave cycles spent waiting."""
    key = PROFILE_KERNEL.get(name)
    if key is None or not PROFILE.exists() or not avg_us:
        return None
    for line in PROFILE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 8 and cells[0].startswith(key):
            try:
                waves, valu, wait = float(cells[1]), float(cells[2]), float(cells[6])
            except ValueError:
                return None
            busy = waves * valu * 4 / (1024 * avg_us * 1e-6 * 2.4e9)
            what = ("VALU issue" if busy >= 0.6 else "load latency" if wait >= 50 else
                    "VALU issue and load latency together")
            return {"bound_by": what, "valu_busy_est": round(busy, 2), "wait_frac_of_wave_cycles": wait / 100,
                    "valu_per_wave": valu, "source": f"{PROFILE.name}: {cells[0]}"}
    return None


def step_model_bytes(nf: int, nd: int) -> float:
    """Compulsory HBM bytes of one WCSPH step per fluid particle: three pair
    passes over the fluid state (40 B each), the writes of v twice, x and
    drho (52 B), hash + sort + ranges (44 B) and the dummy state read by the
    three passes (108 B per dummy, dummies hashed once)."""
    return 3 * 40 + 52 + 44 + 108 * (nd / nf)


def dem_model_bytes(p: float, w: float, b: float) -> float:
    """Compulsory HBM bytes of one DEM ball-step for p ball-pair, w
    ball-triangle and b bond entries per ball: ball state 60 B, pair ids and
    two-way history 104 B per pair, wall contacts 100 B, bonds 112 B, all
    read in each of the two evaluations, the v, w and x writes (72 B), and
    the rebuild every 5th step (44 + 104 p)."""
    return 2 * (60 + 104 * p + 100 * w + 112 * b) + 72 + (44 + 104 * p) / 5


def dem_roofline(g, n: int, ms_per_step: float) -> dict:
    p = g.get("objectPointed").size / n
    w = int((g.get("triCancel") == 0).sum()) / n if g.sizes()[3] > 0 else 0.0
    b = g.get("bondPointed").size / n
    per = dem_model_bytes(p, w, b)
    achieved = per * n / (ms_per_step * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": None,
            "model": "whole step, compulsory bytes per ball-step 2(60 + 104p + 100w + 112b) + 72 + (44 + 104p)/5",
            "p": round(p, 3), "w": round(w, 3), "b": round(b, 3), "bytes_per_ball_step": round(per, 1)}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def algorithmic_flops(kernel: str, pairs_ff: int, pairs_fd: int) -> float:
    """FP32 flops of a pair pass (add/mul/sub = 1, fma = 2, div/sqrt = 1):
    momentum ~100 per pair, continuity ~75 per pair."""
    if kernel.startswith("sph_momentum"):
        return 100.0 * (pairs_ff + pairs_fd)
    if kernel == "sph_continuity":
        return 75.0 * (pairs_ff + pairs_fd)
    return 0.0


def timed_oracle(step_once, budget_s: float, max_steps: int):
    """Steps until budget_s seconds or max_steps steps; returns (steps, s)."""
    k, t1 = 0, time.perf_counter()
    while True:
        step_once()
        k += 1
        el = time.perf_counter() - t1
        if el >= budget_s or k >= max_steps:
            return k, el


def cpu_baseline(spec, budget_s: float = 10.0) -> dict:
    """The FP64 oracle on the host cores: 10 warm-up steps, then 100 steps or
    budget_s seconds, whichever comes first, on every core; then budget_s
    seconds on one thread from the same state.  Setup and I/O excluded."""
    import oracle

    t0 = time.perf_counter()
    o = oracle.OracleSPH(spec)
    L = oracle.lib()
    threads = int(L.o_num_threads())
    for _ in range(10):
        o.step(spec.dt, 1)
    t1 = time.perf_counter()
    steps, el = timed_oracle(lambda: o.step(spec.dt, 1), budget_s, 100)
    L.o_set_num_threads(1)
    s1, e1 = timed_oracle(lambda: o.step(spec.dt, 1), budget_s, 100)
    L.o_set_num_threads(threads)
    total = 10 + steps + s1
    # The GPU path over the same steps from the same state: error against
    # the FP64 oracle at full size.
    import this_studio_amd as ts

    g = ts.SPHScene()
    g.set_eos(spec.eos, spec.gamma)
    g.add_fluid(spec.fluid_pos, None, spec.c0, spec.dx, spec.rho0, spec.nu)
    g.add_dummy(spec.dummy_pos, None, spec.c0, spec.dx, spec.rho0, spec.nu)
    g.initialize(spec.min_bound, spec.max_bound, spec.gravity)
    g.step(spec.dt, total)
    parity = {"steps": total,
              "max_dx_over_spacing": float(abs(g.get("position") - o.get("pos")).max() / spec.dx),
              "max_dv_m_per_s": float(abs(g.get("velocity") - o.get("vel")).max()),
              "max_drho_kg_per_m3": float(abs(g.get("density") - o.get("rho")).max())}
    g.close()
    o.close()
    return {"value": spec.nf * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "logical_cpus": os.cpu_count(),
            "sample": f"FP64 oracle (C, OpenMP, {threads} threads) on the same {spec.nf}-fluid dam break: "
                      f"{steps} steps ({el:.1f} s) after 10 warm-up steps; setup {t1 - t0:.1f} s excluded",
            "single_thread": {"value": spec.nf * s1 / e1, "unit": "particle-steps/s", "cores": 1,
                              "sample": f"the next {s1} steps on one thread ({e1:.1f} s)"},
            "gpu_vs_oracle": parity}


DEM_NULL_STREAM = False


class _NullStream:
    cuda_stream = None


def dem_stream():
    """A HIP stream (default priority) for a DEM scene, which keeps it while
    it lives; the null stream (no graphs) with --dem-null-stream."""
    if DEM_NULL_STREAM:
        return _NullStream()
    import torch

    return torch.cuda.Stream()


def dem_collapse(steps: int, cpu_budget: float, with_cpu: bool) -> dict:
    """The DEM granular collapse (4,096 Hertz-Mindlin spheres in a 0.3 m box
    of 12 triangles, dt = 1e-5 s): ball-steps per second over the first
    `steps` steps after solve() setup (10,000: the whole configured run, the
    spheres landing on the floor after ~3,200), state resident; the same
    scene through the FP64 oracle beside it."""
    import this_studio_amd as ts
    from this_studio_amd import scenes

    s = scenes.granular_collapse()
    # a stream of its own, as a served scene has: whole rebuild cycles replay
    # as graphs (the null stream cannot be captured)
    st = dem_stream()
    g = ts.DEMSceneGPU(s, stream=st.cuda_stream)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    el = time.perf_counter() - t0
    n = g.sizes()[0]
    out = {"workload": "dem_granular_collapse", "balls": n, "triangles": g.sizes()[3], "steps": steps,
           "graph_captures": g.graph_captures(),
           "value": round(n * steps / el, 1), "unit": "ball-steps/s", "ms_per_step": round(el / steps * 1e3, 4),
           "precision": "fp32 contact forces, fp64 positions/velocities/springs",
           "roofline": dem_roofline(g, n, el / steps * 1e3)}
    g.close()  # before its stream goes
    if with_cpu:
        import oracle

        o = oracle.OracleDEM(s)
        L = oracle.lib()
        threads = int(L.o_num_threads())
        o.step(10)
        k, ce = timed_oracle(lambda: o.step(10), cpu_budget, 200)
        k *= 10
        L.o_set_num_threads(1)
        k1, c1 = timed_oracle(lambda: o.step(10), cpu_budget, 200)
        k1 *= 10
        L.o_set_num_threads(threads)
        out["cpu_baseline"] = {"value": round(n * k / ce, 1), "unit": "ball-steps/s", "cores": threads,
                               "kind": "port", "cpu": cpu_model(),
                               "sample": f"FP64 oracle (C, OpenMP, {threads} threads) on the same collapse, "
                                         f"{k} steps ({ce:.1f} s) after 10 warm-up steps",
                               "single_thread": {"value": round(n * k1 / c1, 1), "unit": "ball-steps/s",
                                                 "cores": 1, "sample": f"the next {k1} steps ({c1:.1f} s)"}}
    return out


def cantilever_clumps(steps: int) -> dict:
    """The bonded cantilever with falling clumps: a 128 x 16 x 16 lattice of
    32,768 bonded 2 mm spheres, clamped at x = 0, with a 10 N tip load ramped by
    the hook, and 20,000 tetrahedral 4-sphere clumps above it (112,768 balls in
    all), dt = 5e-7 s.  Ball-steps per second over the first `steps` steps
    after solve() setup (bonds from the initial pairs; 10,000 steps: the whole
    load ramp), state resident."""
    import this_studio_amd as ts
    from this_studio_amd import scenes

    s = scenes.cantilever_with_clumps()
    st = dem_stream()
    g = ts.DEMSceneGPU(s, stream=st.cuda_stream)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    el = time.perf_counter() - t0
    n, nc = g.sizes()[0], g.sizes()[1]
    nbond = int(g.get("bondPointed").size)
    tip = float(g.get("position")[:, 2].min())
    roof = dem_roofline(g, n, el / steps * 1e3)
    g.close()
    return {"workload": "dem_cantilever_clumps", "balls": n, "clumps": nc, "bonds": nbond, "steps": steps,
            "value": round(n * steps / el, 1), "unit": "ball-steps/s", "ms_per_step": round(el / steps * 1e3, 4),
            "lowest_ball_z_m": tip, "roofline": roof,
            "precision": "fp32 contact and bond forces, fp64 positions/velocities/springs/bond state"}


def ice_cylinder_coupled(steps: int) -> dict:
    """The coupled ice sheet on a current: ~980k WCSPH fluid particles at dx =
    4 mm moving at 0.5 m/s in a 2.0 x 0.8 m tank, ~50k bonded ice balls (r =
    4 mm, rho = 917) in two layers on the surface, a 2,048-facet cylinder of R
    = 0.1 m (DEM wall, and dummy particles for the fluid), two-way coupling,
    dt = 5e-6 s.  Particle-steps (fluid + balls) per second over `steps`
    coupled steps after setup and 5 warm-up steps, state resident."""
    import this_studio_amd as ts
    from this_studio_amd import scenes, serving

    spec, d = scenes.ice_cylinder_coupled()
    f = serving.staged_sph(spec)
    f.initialize(spec.min_bound, spec.max_bound, spec.gravity)
    g = ts.DEMSceneGPU(d)
    c = ts.CoupledScene(f, g)
    c.step(spec.dt, 5, spec.gravity)
    f.synchronize()
    t0 = time.perf_counter()
    c.step(spec.dt, steps, spec.gravity)
    f.synchronize()
    el = time.perf_counter() - t0
    nf, nd = f.sizes()[0], f.sizes()[1]
    nb, ntri = g.sizes()[0], g.sizes()[3]
    bonds = g.get("bondIsBonded")
    out = {"workload": "ice_cylinder_coupled", "fluid_particles": nf, "dummy_particles": nd, "ice_balls": nb,
           "bonds": int(bonds.size), "bonds_intact": int(bonds.sum()), "triangles": ntri, "steps": steps,
           "dt": spec.dt, "value": round((nf + nb) * steps / el, 1), "unit": "particle-steps/s (fluid + balls)",
           "ms_per_step": round(el / steps * 1e3, 4),
           "precision": "fp32 pair and contact forces, fp64 positions/velocities/springs"}
    # the fluid half against the WCSPH step model, the ice half against the DEM one
    fluid_b = step_model_bytes(nf, nd) * nf
    dem_b = dem_roofline(g, nb, 1.0)["bytes_per_ball_step"] * nb
    achieved = (fluid_b + dem_b) / (el / steps) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": None,
                       "model": "WCSPH bytes per fluid particle-step + DEM bytes per ball-step, whole coupled step"}
    c.close()
    g.close()
    f.close()
    return out


def serving_mix(count: int, scale: float, slots: int, autotune: str | None = None,
                runs_wanted=("plain", "batched", "batched_autotuned"), hash_file: str | None = None,
                block: int = 100) -> dict:
    """The batched-serving job stream (256 scenes of 2k-200k particles,
    log-uniform; 50/30/20 DEM/WCSPH/coupled; 1k-20k steps; Poisson arrivals
    of mean 50 ms) scaled to finish in seconds: `count` scenes, step counts
    and arrival gaps both times `scale` (so the stream stays as work-bound
    as the full one).  Served twice from the same seed: plain (one slot, one
    launch per stage from the host, every job alone) and batched (`slots`
    concurrent jobs, WCSPH steps as graph replays).  Plain runs every job
    alone, so equal results in the batched run show that the mix changes
    nothing.  With `autotune`, a third run serves the batched path with
    every fluid scene autotuned on admission (the tuning inside the wall
    time; "exact" keeps results bit-identical).  runs_wanted picks runs (the
    full-size stream takes minutes per run); each job's final positions are
    also hashed (SHA-256), and hash_file keeps the hashes so that runs made
    in separate processes can be compared."""
    import hashlib

    import numpy as np

    import this_studio_amd as ts
    from this_studio_amd import scenes, serving

    lo, hi = max(1, int(round(1000 * scale))), max(2, int(round(20000 * scale)))
    jobs = scenes.serving_mix(count, steps_range=(lo, hi), mean_interarrival=0.05 * scale)
    out = {"scenes": count, "steps_range": [lo, hi], "mean_interarrival_s": 0.05 * scale,
           "kinds": {k: sum(j.kind == k for j in jobs) for k in ("dem", "wcsph", "coupled")},
           "memory_budget_GB": 288}
    results = {}
    runs = [("plain", 1, False, None), ("batched", slots, True, None)]
    if autotune:
        runs.append(("batched_autotuned", slots, True, autotune))
        ts.SPHScene.clear_autotune_cache()
    runs = [r for r in runs if r[0] in runs_wanted]
    hashes = {}
    for name, sl, graphs, tune in runs:
        print(f"[bench] serving mix: {name} ({sl} slot(s), graphs={graphs}, autotune={tune}) ...", file=sys.stderr,
              flush=True)
        every = max(1, count // 16)

        def progress(done, total, t, name=name):
            if done % every == 0 or done == total:
                print(f"[bench] serving mix: {name}: {done}/{total} jobs finished, {t:.1f} s", file=sys.stderr,
                      flush=True)

        srv, ids = serving.serve(jobs, slots=sl, graphs=graphs, autotune=tune, progress=progress, block=block)
        st = srv.stats()
        out[name] = {"slots": sl, "graphs": graphs, "autotune": tune, "block": block, "wall_s": round(st["wall"], 4),
                     "particle_steps_per_s": round(st["particleStepsPerSecond"], 1),
                     "p50_step_ms": round(st["p50"] * 1e3, 4), "p99_step_ms": round(st["p99"] * 1e3, 4),
                     "jobs_done": st["jobsDone"], "jobs_failed": st["jobsFailed"], "peak_active": st["peakActive"],
                     "peak_admitted_GB": round(st["peakBytes"] / 1e9, 3)}
        errs = [f"{j.kind} {j.n}: {srv.info(i)['error']}" for j, i in zip(jobs, ids) if srv.info(i)["status"]]
        if errs:
            out[name]["errors"] = errs[:5]
        results[name] = [(srv.result(i, "position") if j.kind != "dem" else srv.result(i, "ballPosition"))
                         if not srv.info(i)["status"] else None for j, i in zip(jobs, ids)]
        hashes[name] = [hashlib.sha256(r.tobytes()).hexdigest() if r is not None else None for r in results[name]]
        srv.close()
    if "plain" in out and "batched" in out:
        out["speedup"] = round(out["batched"]["particle_steps_per_s"] / out["plain"]["particle_steps_per_s"], 3)
    if "plain" in out and "batched_autotuned" in out:
        out["speedup_autotuned"] = round(out["batched_autotuned"]["particle_steps_per_s"]
                                         / out["plain"]["particle_steps_per_s"], 3)
    if len(results) > 1:
        ref = next(iter(results.values()))
        out["results_equal_alone_and_mixed"] = all(
            all(r[k] is not None and np.array_equal(r[k], ref[k]) for r in results.values())
            for k in range(len(jobs)))
    if hash_file:
        Path(hash_file).write_text(json.dumps({"scenes": count, "scale": scale, "runs": hashes,
                                               "summary": {k: out[k] for k in hashes}}))
    return out


def full_dam_break(spec) -> dict:
    """The whole configured run: the dam break from rest to t = 0.5 s
    (ceil(0.5 / dt) steps) on a fresh scene, as served (graph replays),
    with the front position and column height at the end."""
    import math

    import this_studio_amd as ts

    steps = int(math.ceil(0.5 / spec.dt))
    g = ts.SPHScene()
    g.set_eos(spec.eos, spec.gamma)
    g.add_fluid(spec.fluid_pos, None, spec.c0, spec.dx, spec.rho0, spec.nu)
    g.add_dummy(spec.dummy_pos, None, spec.c0, spec.dx, spec.rho0, spec.nu)
    g.initialize(spec.min_bound, spec.max_bound, spec.gravity)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(spec.dt, steps)
    g.synchronize()
    el = time.perf_counter() - t0
    p = g.get("position")
    out = {"t_end_s": steps * spec.dt, "steps": steps, "wall_s": round(el, 3),
           "ms_per_step": round(el / steps * 1e3, 4), "value": round(spec.nf * steps / el, 1),
           "unit": "particle-steps/s", "front_x_m": float(p[:, 0].max()), "max_z_m": float(p[:, 2].max())}
    g.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dx", type=float, default=0.005)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--dem-steps", type=int, default=10000, help="steps of the DEM collapse leg (0: skip it)")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of each CPU-baseline sample (all cores, then one thread)")
    ap.add_argument("--full-run", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the whole 0.5 s dam break from rest")
    ap.add_argument("--mix-scenes", type=int, default=48, help="scenes of the serving-mix leg (0: skip it)")
    ap.add_argument("--beam-steps", type=int, default=10000,
                    help="steps of the cantilever-with-clumps leg (0: skip it)")
    ap.add_argument("--ice-steps", type=int, default=200, help="steps of the coupled ice-cylinder leg (0: skip it)")
    ap.add_argument("--mix-scale", type=float, default=0.05, help="step-count and arrival-gap scale of the mix")
    ap.add_argument("--mix-slots", type=int, default=8, help="concurrent jobs of the batched server")
    ap.add_argument("--mix-block", type=int, default=100, help="steps per step call of a served job")
    ap.add_argument("--mix-runs", default="plain,batched,batched_autotuned",
                    help="comma-separated serving runs to make (plain, batched, batched_autotuned)")
    ap.add_argument("--mix-hashes", default=None, metavar="FILE",
                    help="write each run's per-job result hashes and summary to FILE")
    ap.add_argument("--mix-autotune", choices=("off", "exact", "pairs"), default="exact",
                    help="also serve the batched path with fluid scenes autotuned on admission")
    ap.add_argument("--no-autotune-report", action="store_true",
                    help="skip the autotuner trials on the dam break after the timed steps")
    ap.add_argument("--dem-null-stream", action="store_true",
                    help="run the DEM legs on the null stream, one launch at a time, as for a kernel trace "
                         "(rocprofv3 has ended with a host-side SIGSEGV inside their cycle-graph launches)")
    ap.add_argument("--no-state-io", action="store_true",
                    help="skip timing the scene state export and import after the timed steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 launches) to calibrate HBM counters")
    args = ap.parse_args()
    global DEM_NULL_STREAM
    DEM_NULL_STREAM = args.dem_null_stream
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                          "launches": 5}))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build measures the single-GPU kernel "
              f"(--gpus {args.gpus} requested; multi-GPU sharding is a design, not a run)", file=sys.stderr)
        return 2

    import this_studio_amd as ts
    from this_studio_amd import scenes

    spec = scenes.dam_break(dx=args.dx)
    s = ts.SPHScene()
    s.set_eos(spec.eos, spec.gamma)
    s.add_fluid(spec.fluid_pos, None, spec.c0, spec.dx, spec.rho0, spec.nu)
    s.add_dummy(spec.dummy_pos, None, spec.c0, spec.dx, spec.rho0, spec.nu)
    s.initialize(spec.min_bound, spec.max_bound, spec.gravity)
    print(f"[bench] wcsph dam break: {spec.nf} fluid, {spec.nd} dummies, dt={spec.dt:.3e}", file=sys.stderr,
          flush=True)
    s.step(spec.dt, args.warmup)
    s.synchronize()
    print(f"[bench] warm-up done ({args.warmup} steps)", file=sys.stderr, flush=True)

    # Throughput: the step as served (two-step graph replays, no timing events).
    s.synchronize()
    t0 = time.perf_counter()
    s.step(spec.dt, args.steps)
    s.synchronize()
    el = time.perf_counter() - t0
    # The same steps launched one kernel at a time from the host.
    s.set_graphs(False)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(spec.dt, args.steps)
    s.synchronize()
    el_eager = time.perf_counter() - t0
    s.set_graphs(True)
    # Per-kernel durations: HIP events around every launch on the scene's
    # stream (eager launches; events are not recorded inside graphs).
    s.enable_timing(True)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(spec.dt, args.steps)
    s.synchronize()
    el_timed = time.perf_counter() - t0
    timing = s.timing()
    s.enable_timing(False)
    print(f"[bench] {args.steps} steps: graphs {el * 1e3 / args.steps:.4f} ms/step, eager "
          f"{el_eager * 1e3 / args.steps:.4f}, eager+events {el_timed * 1e3 / args.steps:.4f}", file=sys.stderr,
          flush=True)
    cnt_ff = s.get("neighborCount")
    cnt_fd = s.get("dummyNeighborCount")
    pairs_ff, pairs_fd = int(cnt_ff.sum()), int(cnt_fd.sum())
    print(f"[bench] timed {args.steps} steps in {el:.3f} s; pairs ff={pairs_ff} fd={pairs_fd}", file=sys.stderr,
          flush=True)

    nf = spec.nf
    value = nf * args.steps / el
    kernels = {}
    for name, (launches, total_ms) in timing.items():
        avg_ms = total_ms / max(launches, 1)
        b = algorithmic_bytes(name, nf, pairs_ff, pairs_fd)
        fl = algorithmic_flops(name, pairs_ff, pairs_fd)
        kernels[name] = {"launches": launches, "avg_us": avg_ms * 1e3,
                         "share": total_ms / (el_timed * 1e3),
                         "GBps": (b / (avg_ms * 1e-3) / 1e9) if b else None,
                         "TFLOPs": (fl / (avg_ms * 1e-3) / 1e12) if fl else None}
        print(f"[bench] {name:22s} {launches:6d} x {avg_ms * 1e3:9.1f} us  "
              f"share {kernels[name]['share'] * 100:5.1f}%  "
              f"{(kernels[name]['GBps'] or 0):8.1f} GB/s  {(kernels[name]['TFLOPs'] or 0):6.2f} TF/s",
              file=sys.stderr, flush=True)
    dom = max(timing, key=lambda k: timing[k][1])
    d_launches, d_total = timing[dom]
    d_avg_s = d_total / d_launches * 1e-3
    d_bytes = algorithmic_bytes(dom, nf, pairs_ff, pairs_fd)
    achieved = d_bytes / d_avg_s / 1e9
    traffic, traffic_src = profiled_traffic(dom)
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": traffic, "traffic_source": traffic_src,
                "avg_launch_us": round(d_avg_s * 1e6, 2), "algorithmic_bytes": int(d_bytes),
                "limiter": profiled_limiter(dom, d_avg_s * 1e6)}
    fl = algorithmic_flops(dom, pairs_ff, pairs_fd)
    if fl:
        roofline["fp32_valu_tflops"] = round(fl / d_avg_s / 1e12, 2)
        roofline["fp32_valu_frac"] = round(fl / d_avg_s / 1e12 / FP32_PEAK_TFLOPS, 4)

    # The autotuner on the same scene, after the timed steps (its trials
    # restore the state): what it picks against the built-in lane rules.
    tune = None
    if not args.no_autotune_report:
        ts.SPHScene.clear_autotune_cache()
        tune = s.autotune(spec.dt, trial_steps=5, mode="pairs", use_cache=False)
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)
        ts.SPHScene.clear_autotune_cache()

    # Eviction cost of the same scene: its state to pageable host memory and
    # back (a serving loop evicting and resuming the job), best of 3 each.
    state_io = None
    if not args.no_state_io:
        t_exp, t_imp = [], []
        for _ in range(3):
            t = time.perf_counter()
            blob = s.export_state()
            t_exp.append(time.perf_counter() - t)
        for _ in range(3):
            t = time.perf_counter()
            s.import_state(blob)
            t_imp.append(time.perf_counter() - t)
        state_io = {"bytes": len(blob), "bytes_per_particle": round(len(blob) / nf, 2),
                    "export_ms": round(min(t_exp) * 1e3, 3), "import_ms": round(min(t_imp) * 1e3, 3),
                    "export_GBps": round(len(blob) / min(t_exp) / 1e9, 2),
                    "import_GBps": round(len(blob) / min(t_imp) / 1e9, 2),
                    "roundtrip_equal": s.export_state() == blob}
        print(f"[bench] state export/import: {state_io}", file=sys.stderr, flush=True)

    # whole-step algorithmic HBM rate (sum over the step's kernels)
    step_bytes = sum(algorithmic_bytes(k, nf, pairs_ff, pairs_fd) * timing[k][0] for k in timing) / args.steps
    # whole step against the compulsory-traffic model of one fluid
    # particle-step, at the served step time
    per_particle = step_model_bytes(nf, spec.nd)
    step_achieved = per_particle * nf / (el / args.steps) / 1e9
    measured = None
    step_traffic = profiled_step_traffic()
    if step_traffic is not None:
        measured = {"bytes": round(step_traffic), "per_fluid_particle": round(step_traffic / nf, 1),
                    "x_model": round(step_traffic / (per_particle * nf), 2),
                    "GBps_at_this_step_time": round(step_traffic / (el / args.steps) / 1e9, 1),
                    "source": f"{PROFILE.name}: HBM read + write per launch x launches per step"}
    step_roofline = {"bound": "hbm", "achieved": round(step_achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(step_achieved / HBM_PEAK_GBPS, 4), "bytes_per_particle_step": round(per_particle, 1),
                     "model": "3*40 + 52 + 44 + 108*Nd/Nf compulsory bytes per fluid particle-step",
                     "measured_traffic_per_step": measured}
    full = None
    if args.full_run:
        full = full_dam_break(spec)
        print(f"[bench] full 0.5 s dam break: {full['steps']} steps, {full['ms_per_step']:.4f} ms/step, "
              f"{full['wall_s']:.2f} s", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(spec, args.cpu_budget)
    dem = None
    if args.dem_steps > 0:
        print("[bench] dem granular collapse ...", file=sys.stderr, flush=True)
        dem = dem_collapse(args.dem_steps, min(args.cpu_budget, 10.0), not args.no_cpu_baseline)
        print(f"[bench] dem collapse: {dem['ms_per_step'] * 1e3:.1f} us/step", file=sys.stderr, flush=True)
    beam = None
    if args.beam_steps > 0:
        print("[bench] bonded cantilever + clumps ...", file=sys.stderr, flush=True)
        beam = cantilever_clumps(args.beam_steps)
        print(f"[bench] cantilever + clumps: {beam['balls']} balls, {beam['bonds']} bonds, "
              f"{beam['ms_per_step'] * 1e3:.1f} us/step", file=sys.stderr, flush=True)
    ice = None
    if args.ice_steps > 0:
        print("[bench] coupled ice cylinder ...", file=sys.stderr, flush=True)
        ice = ice_cylinder_coupled(args.ice_steps)
        print(f"[bench] ice cylinder: {ice['fluid_particles']} fluid + {ice['ice_balls']} balls, "
              f"{ice['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    mix = None
    if args.mix_scenes > 0:
        mix = serving_mix(args.mix_scenes, args.mix_scale, args.mix_slots,
                          None if args.mix_autotune == "off" else args.mix_autotune,
                          tuple(args.mix_runs.split(",")), args.mix_hashes, args.mix_block)
    if mix is not None and "speedup" in mix:
        print(f"[bench] serving mix: plain {mix['plain']['particle_steps_per_s']:.3e} particle-steps/s "
              f"p50 {mix['plain']['p50_step_ms']:.3f} ms p99 {mix['plain']['p99_step_ms']:.3f} ms | batched "
              f"{mix['batched']['particle_steps_per_s']:.3e} p50 {mix['batched']['p50_step_ms']:.3f} ms p99 "
              f"{mix['batched']['p99_step_ms']:.3f} ms | x{mix['speedup']:.2f}, equal={mix['results_equal_alone_and_mixed']}",
              file=sys.stderr, flush=True)
        if "batched_autotuned" in mix:
            bt = mix["batched_autotuned"]
            print(f"[bench] serving mix, autotuned: {bt['particle_steps_per_s']:.3e} particle-steps/s p50 "
                  f"{bt['p50_step_ms']:.3f} ms p99 {bt['p99_step_ms']:.3f} ms | x{mix['speedup_autotuned']:.2f}",
                  file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "eager_ms_per_step": round(el_eager / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "wcsph_dam_break", "fluid_particles": nf, "dummy_particles": spec.nd,
                   "dx": spec.dx, "h": 1.3 * spec.dx, "eos": "tait_gamma7", "c0": round(spec.c0, 4),
                   "dt": spec.dt, "pairs_ff": pairs_ff, "pairs_fd": pairs_fd,
                   "step_hbm_GBps": round(step_bytes / (el / args.steps) / 1e9, 1),
                   "precision": "fp32 pair arithmetic (pair distances from fp32 offsets to cell origins), fp64 position/velocity integration, fp32 drho"},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "full_run": full,
        "cpu_baseline": cpu,
        "dem_collapse": dem,
        "serving_mix": mix,
        "autotune": tune,
        "state_io": state_io,
        "cantilever_clumps": beam,
        "ice_cylinder_coupled": ice,
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
