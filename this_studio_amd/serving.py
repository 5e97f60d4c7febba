"""Staging of the batched-serving job stream (scenes.serving_mix) onto a
Server: each job becomes a staged, not yet initialised scene that the server
initialises on admission."""
from __future__ import annotations

import numpy as np

from . import DEMSceneGPU, SPHScene, Server
from . import scenes


def staged_sph(spec: scenes.SPHSceneSpec) -> SPHScene:
    s = SPHScene()
    s.set_eos(spec.eos, spec.gamma)
    s.add_fluid(spec.fluid_pos, spec.fluid_vel, *spec.fluid_props())
    if spec.nd:
        s.add_dummy(spec.dummy_pos, spec.dummy_vel, spec.c0, spec.dx, spec.rho0, spec.nu)
    return s


def submit(server: Server, job: scenes.MixJob, steps: int | None = None, arrival: float | None = None) -> int:
    """Builds the job's scene(s) on the host and submits them; returns the
    job id.  steps / arrival override the job's."""
    steps = job.steps if steps is None else steps
    arrival = job.arrival if arrival is None else arrival
    if job.kind == "wcsph":
        spec = scenes.mix_wcsph(job.n)
        return server.submit(staged_sph(spec), None, spec.gravity, spec.dt, steps, arrival,
                             fluid_bounds=(spec.min_bound, spec.max_bound))
    if job.kind == "dem":
        d = scenes.mix_dem(job.n, job.seed)
        return server.submit(None, DEMSceneGPU(d, initialize=False), d.gravity, d.dt, steps, arrival,
                             dem_bounds=(d.min_bound, d.max_bound))
    if job.kind == "coupled":
        spec, d = scenes.mix_coupled(job.n, job.seed)
        return server.submit(staged_sph(spec), DEMSceneGPU(d, initialize=False), spec.gravity, spec.dt, steps,
                             arrival, fluid_bounds=(spec.min_bound, spec.max_bound),
                             dem_bounds=(d.min_bound, d.max_bound))
    raise ValueError(f"unknown job kind {job.kind!r}")


def serve(jobs: list[scenes.MixJob], slots: int, graphs: bool, memory_budget: float = 288e9,
          block: int = 100, autotune: str | None = None, progress=None,
          tune_min_steps: int = 2000) -> tuple[Server, list[int]]:
    """Submits every job to a new server and runs it; returns the server
    (for results) and the job ids in submission order.  progress: see
    Server.set_progress."""
    srv = Server(slots=slots, memory_budget=memory_budget, graphs=graphs, block=block, autotune=autotune,
                 tune_min_steps=tune_min_steps)
    ids = [submit(srv, j) for j in jobs]
    if progress is not None:
        srv.set_progress(progress)
    srv.run()
    return srv, ids


def all_latencies(srv: Server, ids: list[int]) -> np.ndarray:
    parts = [srv.latencies(i) for i in ids]
    return np.concatenate(parts) if parts else np.zeros(0)
