"""Continuous batching: jobs admitted while others run give the same result alone and in the mix,
and the same result as the plain single-system step (bitwise: every per-system quantity is computed
independently of the other resident systems)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
HYP = (11, 6, 4)


def _classes(gpu):
    import nepx
    net = nepx.EnergyNet.random(40, [64, 64], seed=4, device=gpu)
    return {"tet10": (nepx.get_shape("tet10"), HYP, net), "cube8": (nepx.get_shape("cube8"), HYP, net)}


def _jobs():
    from nepx.configs import make_system
    from nepx.serve import Job
    specs = [("tet10", make_system("tet10", 300, 0.1, kT=0.8, seed=11), 30),
             ("tet10", make_system("tet10", 700, 0.3, kT=1.4, seed=12), 17),
             ("cube8", make_system("cube8", 216, 0.25, kT=1.0, seed=13, non_overlap=False), 23),
             ("tet10", make_system("tet10", 450, 0.2, kT=0.6, seed=14), 41),
             ("cube8", make_system("cube8", 343, 0.3, kT=1.9, seed=15, non_overlap=False), 12)]
    return [Job(k, c, s, n, 0.0) for k, (c, s, n) in enumerate(specs)]


def test_served_mix_equals_alone_and_plain(gpu):
    from nepx.serve import Server, run_plain
    classes = _classes(gpu)
    # 1280 particles (5 blocks of 256) per class: the third tet job waits until one of the first two retires
    srv = Server(classes, capacity=1280, max_systems=4, chunk=5, device=gpu)
    stats = srv.run(_jobs())
    mix = {j.jid: j for j in _finished(srv, stats)}
    assert stats["jobs_finished"] == 5 and stats["jobs_rejected"] == 0
    waited = [j for j in srv.last_finished if j.admitted_at > 0]
    assert waited, "no job was admitted while others were running"
    for job in _jobs():
        alone_srv = Server(classes, capacity=1280, max_systems=4, chunk=7, device=gpu)
        alone_srv.run([job])
        a = job.result
        b = mix[job.jid].result
        assert a["thermo"]["steps"] == job.steps == b["thermo"]["steps"]
        for key in "xvqp":
            assert np.array_equal(a[key], b[key]), (job.jid, key, float(np.abs(a[key] - b[key]).max()))
    # the plain path (single-system arena, device init) reaches the same states
    jobs = _jobs()
    st = run_plain(jobs, classes, chunk=6, device=gpu)
    assert st["jobs_finished"] == 5
    for job in jobs:
        for key in "xvqp":
            assert np.array_equal(job.result[key], mix[job.jid].result[key]), (job.jid, key)


def _finished(srv, stats):
    return srv.last_finished


def test_oversized_job_is_rejected_not_queued(gpu):
    from nepx.configs import make_system
    from nepx.serve import Job, Server
    srv = Server(_classes(gpu), capacity=512, max_systems=2, chunk=5, device=gpu)
    big = Job(0, "tet10", make_system("tet10", 700, 0.3, seed=1), 10, 0.0)
    small = Job(1, "tet10", make_system("tet10", 200, 0.05, seed=2), 10, 0.0)
    st = srv.run([big, small])
    assert st["jobs_rejected"] == 1 and st["jobs_finished"] == 1 and "error" in big.result
