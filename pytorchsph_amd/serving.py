"""Serving: a request loop with continuous batching over the batched engine.

Jobs (a scene and a number of steps) arrive over time.  The server keeps
one running batch (a MixedBatch: delta-SPH and DFSPH parts on their own
streams) whose particle state stays on the device between steps.  Every
tick advances every job in the batch by one step (one graph replay per
part).  The batch is rebuilt -- "repacked" -- when

  * jobs are waiting and the HBM budget and the particle cap admit the next
    one (at most once every `min_repack_s`), or
  * finished jobs hold more than `waste_frac` of the batch's particles, or
  * every job in the batch has finished.

A repack exports each unfinished job's state exactly (positions,
velocities, density, particle ids and the pending cell order, see
SceneBatch.export) and builds the new batch from those and the newly
admitted jobs, so a job advances exactly as it would alone: its kernels
read only its own scene's data.  Admission is by device footprint
(SceneBatch.footprint, the bytes a batch of the job allocates) against
`budget_bytes` (default 85 % of the free device memory at start).

A job's result is read back the tick it finishes; a finished job keeps
being stepped (and ignored) until the next repack drops it.  A job whose
scene goes NaN, leaves its cell table or has its pair list truncated
finishes at once, failed, with that status (`fail_fast`).  A job's status
is the OR of every batch it ran in.  Per-step latency is the wall time of
the tick in which a job advanced.

The particle cap (`max_particles`, default 2M) bounds the tick.  A tick
lasts as long as its slowest scene: the DFSPH iterations of a tick are the
maximum over its scenes, and the early ones run over nearly every row, so
a batch of more scenes costs more per particle.  Below ~1.5M particles the
per-tick overheads take over instead.  On the 10 s Poisson stream of the
benchmark, caps of 1, 1.5, 2, 3, 6 and 12M particles gave 2.8, 3.1, 3.1,
2.5, 1.6 and 1.7e8 completed particle-steps/s with p99 ticks of 11, 15,
19, 59, 92 and 128 ms (more jobs wait instead).
"""
from __future__ import annotations

import os
import time
from collections import deque
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import ST_NAN, ST_OUTSIDE, ST_OVERFLOW
from .engine import MixedBatch, SceneBatch
from .export import write_job_result
from .scenes import Scene


@dataclass
class Job:
    id: int
    scene: Scene
    steps: int
    t_arrive: float = 0.0          # seconds after the start of the loop
    bytes: int = 0
    steps_done: int = 0
    t_admit: float | None = None
    t_done: float | None = None
    result: dict | None = None     # x, v, rho in the original particle order, status
    carry: dict | None = None      # exported state between repacks
    failed: bool = False           # stopped early: NaN, outside its table or truncated pairs
    status: int = 0                # status bits raised in the batches it has left

    @property
    def n(self) -> int:
        return self.scene.n

    @property
    def finished(self) -> bool:
        return self.failed or self.steps_done >= self.steps


class Server:
    def __init__(self, device="cuda", budget_bytes: int | None = None,
                 max_particles: int = 2_000_000, min_repack_s: float = 0.5,
                 waste_frac: float = 0.5, log=None, export_dir: str | None = None,
                 fail_fast: bool = True, **batch_kw):
        self.device = device
        # finish a job as failed on SPH_ST_NAN / SPH_ST_OUTSIDE /
        # SPH_ST_OVERFLOW (one status read per tick)
        self.fail_fast = bool(fail_fast)
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            budget_bytes = int(0.85 * free)
        self.budget = int(budget_bytes)
        self.max_particles = int(max_particles)
        self.min_repack_s = float(min_repack_s)
        self.waste_frac = float(waste_frac)
        self.batch_kw = batch_kw
        self.log = log
        # finished jobs are written as export_dir/job_<id>.hdf5 in the
        # reference's export layout (export.write_job_result)
        self.export_dir = export_dir
        self.batch: MixedBatch | None = None
        self.live: list[Job] = []
        self.repacks = 0
        self.rejected: list[Job] = []
        self.repack_times: list[dict] = []      # seconds per phase of each repack
        self.latency: list[float] = []          # one sample per (job, step)
        self.ticks: list[tuple[float, int]] = []  # (seconds, particles advanced)

    # ------------------------------------------------------------ helpers
    def footprint(self, job: Job) -> int:
        if not job.bytes:
            job.bytes = SceneBatch.footprint([job.scene], **self.batch_kw)
        return job.bytes

    def _finish(self, job: Job, k: int, now: float):
        st = self.batch.scene_state(k, extra=self.export_dir is not None)
        b, pos = self.batch.index[k]
        st["status"] = int(b.scene_status(pos)) | job.status
        job.result = st
        job.t_done = now
        if self.export_dir is not None:
            os.makedirs(self.export_dir, exist_ok=True)
            path = os.path.join(self.export_dir, f"job_{job.id}.hdf5")
            write_job_result(path, job.scene, st, job.steps_done)
            st["export"] = path

    def _advance(self, pos_of: dict, seconds: float, now: float):
        """Book-keeping after one step of the whole batch.  A job whose
        scene raised SPH_ST_NAN (the DFSPH pressure went NaN / Inf, where the
        reference raises, dfsph.py:179-180), SPH_ST_OUTSIDE (a particle
        left the scene's cell table, so the pair lists no longer hold all
        its pairs) or SPH_ST_OVERFLOW (the batch's shared pair capacity was
        exceeded and this scene's rows were truncated) is finished at once
        with that status, failed."""
        moved = 0
        status = {id(b): b.status() for b in self.batch.parts} if self.fail_fast else {}
        for j in self.live:
            if j.finished:
                continue
            j.steps_done += 1
            moved += j.n
            self.latency.append(seconds)
            if self.fail_fast:
                b, pos = self.batch.index[pos_of[j.id]]
                if (int(status[id(b)][pos]) | j.status) & (ST_NAN | ST_OUTSIDE | ST_OVERFLOW):
                    j.failed = True
            if j.finished:
                self._finish(j, pos_of[j.id], now)
        self.ticks.append((seconds, moved))

    def _repack(self, queue: deque, now: float, clock, start):
        t_begin = clock()
        keep = [j for j in self.live if not j.finished]
        for k, j in enumerate(self.live):
            if not j.finished:
                j.scene, j.carry = self.batch.export(k)
                j.status |= int(j.carry.get("status", 0))
        self.batch = None
        t_export = clock()
        used = sum(self.footprint(j) for j in keep)
        parts = sum(j.n for j in keep)
        admitted = []
        while queue:
            j = queue[0]
            b = self.footprint(j)
            if used + b > self.budget or (parts + j.n > self.max_particles and (keep or admitted)):
                break
            queue.popleft()
            used += b
            parts += j.n
            j.t_admit = now
            admitted.append(j)
        if not keep and not admitted and queue:
            # larger than the whole budget: it can never run
            j = queue.popleft()
            j.result = dict(rejected=True)
            j.t_done = now
            self.rejected.append(j)
        self.live = keep + admitted
        self.repacks += 1
        if not self.live:
            return
        t0 = clock()
        self.batch = MixedBatch([j.scene for j in self.live], self.device,
                                carry=[j.carry for j in self.live], **self.batch_kw)
        torch.cuda.synchronize()
        t_build = clock()
        # record the graphs, then the tick's step is an ordinary replay
        cap = self.batch.capture(warm=False)
        t_step = clock()
        self.batch.step()
        torch.cuda.synchronize()
        dt = clock() - t0
        self.repack_times.append(dict(export=t_export - t_begin, admit=t0 - t_export,
                                      build=t_build - t0, record=cap["record"],
                                      step=clock() - t_step))
        for j in self.live:
            j.carry = None
        self._advance({j.id: k for k, j in enumerate(self.live)}, dt, clock() - start)
        if self.log:
            rt = self.repack_times[-1]
            its = [b.dfsph_iterations() for b in self.batch.parts if b.scheme == "dfsph"]
            it_s = (f", DFSPH iterations max {int(its[0].max())} mean {its[0].mean():.0f}"
                    if its and its[0].size else "")
            self.log(f"repack {self.repacks}: {len(keep)} kept + {len(admitted)} admitted, "
                     f"{parts} particles, {used / 2**30:.2f} GiB, {dt * 1e3:.0f} ms ("
                     + ", ".join(f"{k} {v * 1e3:.0f}" for k, v in rt.items()) + ")" + it_s)

    # --------------------------------------------------------------- loop
    def run(self, jobs: list[Job], time_limit: float | None = None, drain: bool = True,
            clock=time.perf_counter) -> dict:
        """Serves `jobs` (arrival times relative to the start).  Stops when
        every job has finished (drain) or at `time_limit` seconds."""
        pending = deque(sorted(jobs, key=lambda j: j.t_arrive))
        queue: deque = deque()
        start = clock()
        last_repack = -1e30
        while True:
            now = clock() - start
            if time_limit is not None and now >= time_limit:
                break
            while pending and pending[0].t_arrive <= now:
                queue.append(pending.popleft())
            live_n = sum(j.n for j in self.live)
            dead_n = sum(j.n for j in self.live if j.finished)
            running = [j for j in self.live if not j.finished]
            can_admit = bool(queue) and (
                sum(self.footprint(j) for j in running) + self.footprint(queue[0]) <= self.budget
                and sum(j.n for j in running) + queue[0].n <= self.max_particles)
            if (self.batch is None and queue) \
                    or (self.batch is not None and dead_n == live_n) \
                    or (self.batch is not None and dead_n > self.waste_frac * live_n) \
                    or (can_admit and now - last_repack >= self.min_repack_s):
                self._repack(queue, now, clock, start)
                last_repack = clock() - start
                continue
            if self.batch is None:
                if not pending and not queue:
                    break
                if not drain and not pending and time_limit is None:
                    break
                wait = (pending[0].t_arrive - now) if pending else 0.001
                time.sleep(max(0.0, min(wait, 0.01)))
                continue
            t0 = clock()
            self.batch.step()
            torch.cuda.synchronize()
            self._advance({j.id: k for k, j in enumerate(self.live)}, clock() - t0,
                          clock() - start)
        wall = clock() - start
        done = [j for j in jobs if j.result is not None and not j.result.get("rejected")]
        lat = np.array(self.latency) if self.latency else np.zeros(1)
        moved = sum(m for _, m in self.ticks)
        return dict(
            wall_s=wall,
            particle_steps=moved,
            particle_steps_per_s=moved / wall if wall > 0 else 0.0,
            p50_step_ms=float(np.percentile(lat, 50) * 1e3),
            p99_step_ms=float(np.percentile(lat, 99) * 1e3),
            jobs=len(jobs), jobs_done=len(done),
            jobs_running=sum(1 for j in self.live if not j.finished),
            jobs_waiting=len(queue) + len(pending),
            repacks=self.repacks,
            jobs_rejected=len(self.rejected),
            # finished jobs whose scene raised a status bit: 1 outside the
            # cell table, 2 pair list truncated, 4 NaN pressure, 8 iteration cap
            jobs_flagged={str(bit): sum(1 for j in done if j.result.get("status", 0) & bit)
                          for bit in (1, 2, 4, 8)},
            jobs_failed=sum(1 for j in done if j.failed),
            mean_queue_s=float(np.mean([j.t_admit - j.t_arrive for j in jobs
                                        if j.t_admit is not None])) if any(
                j.t_admit is not None for j in jobs) else None,
        )
