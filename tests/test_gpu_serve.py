"""GPU tests of the batched scene server (demsph_server_*): every job's
result is bit-identical alone (plain path: one slot, no graphs) and in a
concurrent mix (several slots, graph replays); served jobs agree with the
FP64 oracle; admission respects the memory budget and arrival times."""
from dataclasses import replace

import numpy as np
import pytest

import oracle
import this_studio_amd as ts
from this_studio_amd import scenes, serving
from this_studio_amd.scenes import MixJob

pytestmark = pytest.mark.gpu

SMALL = [
    MixJob("wcsph", 3000, 80, 0.0, 11),
    MixJob("dem", 3000, 80, 0.0, 12),
    MixJob("coupled", 4000, 60, 0.0, 13),
    MixJob("wcsph", 8000, 81, 0.0, 14),
    MixJob("dem", 9000, 70, 0.0, 15),
    MixJob("coupled", 3000, 50, 0.0, 16),
]


def _fields(kind):
    return {"wcsph": ("position", "velocity"), "dem": ("ballPosition", "ballVelocity"),
            "coupled": ("position", "velocity", "ballPosition", "ballVelocity")}[kind]


def test_mix_results_equal_alone():
    print("[serve] mix of", len(SMALL), "jobs, 3 slots, graphs", flush=True)
    srv, ids = serving.serve(SMALL, slots=3, graphs=True, block=32)
    for job, i in zip(SMALL, ids):
        info = srv.info(i)
        assert info["status"] == 0, info["error"]
        assert info["steps_done"] == job.steps
        lat = srv.latencies(i)
        assert lat.size == job.steps - 1
        assert (lat > 0).all()
    st = srv.stats()
    assert st["jobsDone"] == len(SMALL) and st["jobsFailed"] == 0
    assert st["peakActive"] == 3  # all jobs arrive at once: every slot takes one
    assert 0 < st["p50"] <= st["p99"]
    for job, i in zip(SMALL, ids):
        print("[serve] alone:", job, flush=True)
        alone, (k,) = serving.serve([job], slots=1, graphs=False)
        for f in _fields(job.kind):
            assert np.array_equal(srv.result(i, f), alone.result(k, f)), (job, f)
        alone.close()


def test_many_short_jobs_capture_while_others_allocate():
    """Many short jobs on 8 slots: graph captures in some slots overlap the
    set-up, read-back and teardown (allocation, release, copies) of others.
    Every job must complete and equal its plain run."""
    kinds = ("dem", "wcsph", "dem", "coupled")
    jobs = [MixJob(kinds[k % 4], 2000 + 300 * k, 24 + (k % 5) * 4, 0.0, 100 + k) for k in range(24)]
    srv, ids = serving.serve(jobs, slots=8, graphs=True, block=8)
    for job, i in zip(jobs, ids):
        assert srv.info(i)["status"] == 0, (job, srv.info(i)["error"])
    plain, pids = serving.serve(jobs, slots=1, graphs=False)
    for job, i, k in zip(jobs, ids, pids):
        for f in _fields(job.kind):
            assert np.array_equal(srv.result(i, f), plain.result(k, f)), (job, f)
    srv.close()
    plain.close()


def test_autotuned_serving_keeps_results():
    """Jobs autotuned on admission ("exact": list variants only) give the
    same results, bit for bit, as the plain untuned path."""
    ts.SPHScene.clear_autotune_cache()
    jobs = [SMALL[0], SMALL[2], SMALL[3]]
    srv, ids = serving.serve(jobs, slots=3, graphs=True, block=32, autotune="exact", tune_min_steps=0)
    for job, i in zip(jobs, ids):
        assert srv.info(i)["status"] == 0, srv.info(i)["error"]
    plain, pids = serving.serve(jobs, slots=1, graphs=False)
    for job, i, k in zip(jobs, ids, pids):
        for f in _fields(job.kind):
            assert np.array_equal(srv.result(i, f), plain.result(k, f)), (job, f)
    srv.close()
    plain.close()
    ts.SPHScene.clear_autotune_cache()


def test_autotune_on_admission_with_eight_slots():
    """Autotuning on admission while 8 slot threads share the hardware
    queues (the trial timings read HIP events then): no job fails, and every
    job equals its plain run bit for bit."""
    ts.SPHScene.clear_autotune_cache()
    kinds = ("wcsph", "dem", "wcsph", "coupled")
    jobs = [MixJob(kinds[k % 4], 2500 + 700 * k, 30 + (k % 3) * 6, 0.0, 300 + k) for k in range(16)]
    print("[serve] 16 jobs, 8 slots, autotune on admission", flush=True)
    srv, ids = serving.serve(jobs, slots=8, graphs=True, block=16, autotune="exact", tune_min_steps=0)
    st = srv.stats()
    assert st["jobsFailed"] == 0 and st["jobsDone"] == len(jobs), [srv.info(i)["error"] for i in ids]
    plain, pids = serving.serve(jobs, slots=1, graphs=False)
    for job, i, k in zip(jobs, ids, pids):
        for f in _fields(job.kind):
            assert np.array_equal(srv.result(i, f), plain.result(k, f)), (job, f)
    srv.close()
    plain.close()
    ts.SPHScene.clear_autotune_cache()


def test_served_jobs_match_oracle():
    """A WCSPH and a DEM job served together against the FP64 oracle."""
    jobs = [MixJob("wcsph", 2000, 40, 0.0, 21), MixJob("dem", 2000, 200, 0.0, 22)]
    srv, ids = serving.serve(jobs, slots=2, graphs=True)
    spec = scenes.mix_wcsph(2000)
    o = oracle.OracleSPH(spec)
    o.step(spec.dt, 40)
    # FP32 pair arithmetic, FP64 integration: within 1e-5 of the spacing
    assert np.abs(srv.result(ids[0], "position") - o.get("pos")).max() / spec.dx < 1e-5
    assert np.abs(srv.result(ids[0], "velocity") - o.get("vel")).max() < 1e-4
    d = scenes.mix_dem(2000, 22)
    od = oracle.OracleDEM(d)
    od.step(200)
    err = np.abs(srv.result(ids[1], "ballPosition") - od.get("pos")).max() / 0.01
    assert err < 1e-5, err


def test_budget_limits_concurrency_and_rejects_oversized_jobs():
    job = MixJob("wcsph", 3000, 30, 0.0, 31)
    probe = ts.Server(slots=1)
    pid = serving.submit(probe, job)
    need = probe.info(pid)["bytes"]
    probe.close()  # destroys the staged scene
    srv = ts.Server(slots=4, memory_budget=1.5 * need, graphs=True)
    ids = [serving.submit(srv, replace(job, seed=31 + k)) for k in range(4)]
    big = serving.submit(srv, MixJob("wcsph", 50000, 10, 0.0, 39))
    srv.run()
    st = srv.stats()
    assert st["peakActive"] == 1  # only one job fits the budget at a time
    assert st["peakBytes"] <= 1.5 * need
    assert st["jobsDone"] == 4 and st["jobsFailed"] == 1
    for i in ids:
        assert srv.info(i)["status"] == 0
    bi = srv.info(big)
    assert bi["status"] == 2 and "budget" in bi["error"]
    # FIFO: admissions in submission order, each after the previous finished
    adm = [srv.info(i)["admitted"] for i in ids]
    fin = [srv.info(i)["finished"] for i in ids]
    assert adm == sorted(adm)
    assert all(adm[k + 1] >= fin[k] for k in range(3))


def test_arrivals_are_respected():
    jobs = [MixJob("dem", 2000, 20, 0.0, 41), MixJob("dem", 2000, 20, 0.25, 42)]
    srv, ids = serving.serve(jobs, slots=2, graphs=True)
    assert srv.info(ids[1])["admitted"] >= 0.25
    assert srv.stats()["wall"] >= 0.25


def test_memory_estimate_covers_allocation():
    """The admission estimate of a WCSPH job against what its scene allocates
    on initialisation and its first steps (device free memory before and
    after)."""
    import torch

    spec = scenes.mix_wcsph(150000)
    srv = ts.Server(slots=1)
    est = srv.info(serving.submit(srv, MixJob("wcsph", 150000, 1, 0.0, 1)))["bytes"]
    srv.close()
    s = serving.staged_sph(spec)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    s.initialize(spec.min_bound, spec.max_bound, spec.gravity)
    s.step(spec.dt, 2)
    s.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    used = free0 - free1
    print(f"[serve] wcsph {spec.nf}: estimate {est / 2**20:.1f} MiB, allocated {used / 2**20:.1f} MiB", flush=True)
    assert est >= used - (16 << 20)
    assert est <= 2 * used + (16 << 20)
    s.close()
