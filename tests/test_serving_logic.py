"""Host logic of the serving loop (admission, repacking, step accounting)
with a stand-in batch that counts steps per scene instead of simulating."""
import dataclasses

import numpy as np
import pytest

from pytorchsph_amd import serving
from pytorchsph_amd.scenes import Scene


def tiny_scene(n):
    z = np.zeros((n, 2), np.float32)
    return Scene(name="t", scheme="dsph", boundary="solid", x=z, v=z, area=1.0, support=1.0,
                 rho0=1.0, bx=np.zeros((0, 2), np.float32), bgx=np.zeros((0, 2), np.float32),
                 bvol=np.zeros(0, np.float32), bterm=None, vmin=(0, 0), vmax=(1, 1))


class CountingBatch:
    """Stands in for MixedBatch: scene k's 'state' is its step count."""
    built = []

    def __init__(self, scenes, device, carry=None, **kw):
        carry = carry or [None] * len(scenes)
        self.scenes = scenes
        self.count = [c["steps"] if c else 0 for c in carry]
        CountingBatch.built.append(sum(s.n for s in scenes))

    def capture(self, warm=True):
        if warm:
            self.step()
        return dict(eager=0.0, record=0.0)

    def step(self):
        self.count = [c + 1 for c in self.count]

    def scene_state(self, k, extra=False):
        return dict(steps=self.count[k])

    @property
    def index(self):
        return {k: (self, k) for k in range(len(self.scenes))}

    # scene k's status bits from step `at` on: {scene name: (at, bits)}
    fail_at: dict = {}

    @property
    def parts(self):
        return [self]

    def status(self):
        return np.array([bits if s.name in self.fail_at and c >= self.fail_at[s.name][0] else 0
                         for s, c in zip(self.scenes, self.count)
                         for bits in [self.fail_at.get(s.name, (0, 0))[1]]], dtype=np.int64)

    def scene_status(self, k):
        return int(self.status()[k])

    def export(self, k):
        # a new batch starts with clear status words, as a device batch
        # would if the carry's status were dropped
        return (dataclasses.replace(self.scenes[k]),
                dict(steps=self.count[k], status=self.scene_status(k)))


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.01
        return self.t


@pytest.fixture
def fake(monkeypatch):
    CountingBatch.built = []
    CountingBatch.fail_at = {}
    monkeypatch.setattr(serving, "MixedBatch", CountingBatch)
    monkeypatch.setattr(serving.torch.cuda, "synchronize", lambda *a: None)
    return CountingBatch


def jobs_of(specs):
    out = []
    for k, (n, steps, t) in enumerate(specs):
        j = serving.Job(id=k, scene=tiny_scene(n), steps=steps, t_arrive=t)
        j.bytes = 100 * n
        out.append(j)
    return out


def test_every_job_gets_exactly_its_steps(fake):
    jobs = jobs_of([(10, 5, 0.0), (20, 30, 0.0), (15, 3, 0.1), (5, 12, 0.3), (8, 1, 0.35)])
    srv = serving.Server(budget_bytes=10**9, min_repack_s=0.05, waste_frac=0.5)
    out = srv.run(jobs, clock=Clock())
    assert out["jobs_done"] == len(jobs)
    for j in jobs:
        assert j.result["steps"] == j.steps, j.id
        assert j.t_admit >= j.t_arrive
    assert out["particle_steps"] == sum(j.n * j.steps for j in jobs)
    assert len(srv.latency) == sum(j.steps for j in jobs)


def test_admission_respects_budget_and_particle_cap(fake):
    jobs = jobs_of([(10, 50, 0.0), (10, 50, 0.0), (10, 50, 0.0), (10, 5, 0.0)])
    # budget for two jobs at a time
    srv = serving.Server(budget_bytes=2 * 100 * 10, min_repack_s=0.0, waste_frac=0.9)
    out = srv.run(jobs, clock=Clock())
    assert out["jobs_done"] == 4
    assert max(fake.built) <= 20
    fake.built = []
    jobs = jobs_of([(10, 20, 0.0), (10, 20, 0.0), (10, 20, 0.0)])
    srv = serving.Server(budget_bytes=10**9, max_particles=25, min_repack_s=0.0)
    srv.run(jobs, clock=Clock())
    assert max(fake.built) <= 25
    assert all(j.result["steps"] == 20 for j in jobs)


def test_job_larger_than_budget_is_rejected(fake):
    jobs = jobs_of([(1000, 5, 0.0), (10, 5, 0.0)])
    srv = serving.Server(budget_bytes=100 * 500, min_repack_s=0.0)
    out = srv.run(jobs, clock=Clock())
    assert jobs[0].result == {"rejected": True}
    assert jobs[1].result["steps"] == 5
    assert out["jobs_rejected"] == 1 and out["jobs_done"] == 1


def test_time_limit_stops_with_backlog(fake):
    jobs = jobs_of([(10, 10**6, 0.0), (10, 10, 5.0)])
    srv = serving.Server(budget_bytes=10**9, min_repack_s=0.0)
    out = srv.run(jobs, time_limit=1.0, clock=Clock())
    assert out["jobs_done"] == 0 and out["jobs_running"] == 1 and out["jobs_waiting"] == 1


def test_footprint_on_meta_device():
    """Admission sizes come from a meta-device build: no device memory and
    no GPU needed, for both schemes, growing with the particle count."""
    from pytorchsph_amd import scenes as S
    from pytorchsph_amd.engine import SceneBatch
    for scheme in ("dsph", "dfsph"):
        small = SceneBatch.footprint([S.scene_for_count("dambreak", scheme, 2000, 1)])
        large = SceneBatch.footprint([S.scene_for_count("dambreak", scheme, 8000, 1)])
        assert 0 < small < large
        assert 200 * 8000 < large < 4000 * 8000


def test_job_whose_scene_fails_is_finished_failed(fake):
    # scene "bad" raises NaN (bit 4) at its third step: it stops there, as
    # the reference raises, and the other jobs run their full counts
    jobs = jobs_of([(10, 20, 0.0), (12, 20, 0.0), (8, 5, 0.0)])
    jobs[1].scene = dataclasses.replace(jobs[1].scene, name="bad")
    fake.fail_at = {"bad": (3, 4)}
    srv = serving.Server(budget_bytes=10**9, min_repack_s=0.0, waste_frac=0.9)
    out = srv.run(jobs, clock=Clock())
    assert jobs[1].failed and jobs[1].steps_done == 3
    assert jobs[1].result["status"] & 4
    assert not jobs[0].failed and jobs[0].result["steps"] == 20
    assert jobs[2].result["steps"] == 5
    assert out["jobs_failed"] == 1 and out["jobs_flagged"]["4"] == 1


def test_fail_fast_off_runs_failed_scenes_to_the_end(fake):
    jobs = jobs_of([(10, 6, 0.0)])
    jobs[0].scene = dataclasses.replace(jobs[0].scene, name="bad")
    fake.fail_at = {"bad": (2, 1)}
    srv = serving.Server(budget_bytes=10**9, min_repack_s=0.0, fail_fast=False)
    out = srv.run(jobs, clock=Clock())
    assert not jobs[0].failed and jobs[0].steps_done == 6
    assert out["jobs_failed"] == 0 and out["jobs_flagged"]["1"] == 1


def test_status_raised_before_a_repack_reaches_the_result(fake):
    """A status bit (here the iteration cap, 8) raised in one batch stays in
    the job's result after repacks into batches whose status words start
    clear."""
    jobs = jobs_of([(10, 40, 0.0), (10, 5, 0.0), (10, 5, 0.2), (10, 5, 0.4)])
    jobs[0].scene = dataclasses.replace(jobs[0].scene, name="capped")
    # raised at the job's second step; sticky in the batch that raised it,
    # clear in every batch built after it
    fake.fail_at = {}

    class Once(CountingBatch):
        def step(self):
            super().step()
            self.raised = getattr(self, "raised", set()) | {
                k for k, (s, c) in enumerate(zip(self.scenes, self.count))
                if s.name == "capped" and c == 2}

        def status(self):
            r = getattr(self, "raised", set())
            return np.array([8 if k in r else 0 for k in range(len(self.scenes))],
                            dtype=np.int64)

    serving.MixedBatch = Once
    srv = serving.Server(budget_bytes=10**9, min_repack_s=0.0, waste_frac=0.5)
    out = srv.run(jobs, clock=Clock())
    assert srv.repacks >= 2
    assert jobs[0].result["steps"] == 40 and not jobs[0].failed
    assert jobs[0].result["status"] & 8
    assert out["jobs_flagged"]["8"] == 1


def test_truncated_pair_list_fails_the_job(fake):
    """SPH_ST_OVERFLOW (2): the job's rows were cut at the batch's pair
    capacity, so it is finished failed like a NaN scene."""
    jobs = jobs_of([(10, 20, 0.0), (12, 20, 0.0)])
    jobs[1].scene = dataclasses.replace(jobs[1].scene, name="bad")
    fake.fail_at = {"bad": (4, 2)}
    srv = serving.Server(budget_bytes=10**9, min_repack_s=0.0)
    out = srv.run(jobs, clock=Clock())
    assert jobs[1].failed and jobs[1].steps_done == 4
    assert jobs[1].result["status"] & 2
    assert jobs[0].result["steps"] == 20
    assert out["jobs_failed"] == 1 and out["jobs_flagged"]["2"] == 1
