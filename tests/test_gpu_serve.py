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
    net60 = nepx.EnergyNet.random(60, [128, 128], seed=5, device=gpu)
    return {"tet10": (nepx.get_shape("tet10"), HYP, net), "cube8": (nepx.get_shape("cube8"), HYP, net),
            "cyl12": (nepx.get_shape("cyl12"), (17, 6, 6), net60)}


def _jobs():
    from nepx.configs import make_system
    from nepx.serve import Job
    specs = [("tet10", make_system("tet10", 300, 0.1, kT=0.8, seed=11), 30),
             ("tet10", make_system("tet10", 700, 0.3, kT=1.4, seed=12), 17),
             ("cube8", make_system("cube8", 216, 0.25, kT=1.0, seed=13, non_overlap=False), 23),
             ("cyl12", make_system("cyl12", 200, 0.1, kT=1.2, seed=16), 19),
             ("tet10", make_system("tet10", 450, 0.2, kT=0.6, seed=14), 41),
             ("cube8", make_system("cube8", 343, 0.3, kT=1.9, seed=15, non_overlap=False), 12),
             ("tet10", make_system("tet10", 8192, 0.25, kT=1.0, seed=17), 9),
             ("cyl12", make_system("cyl12", 343, 0.2, kT=0.7, seed=18), 14)]
    return [Job(k, c, s, n, 0.0) for k, (c, s, n) in enumerate(specs)]


CAP = 9472   # 37 blocks of 256 per class: the 8192-particle tet job waits until a smaller one retires


def test_served_mix_equals_alone_and_plain(gpu):
    from nepx.serve import Server, run_plain
    classes = _classes(gpu)
    srv = Server(classes, capacity=CAP, max_systems=4, chunk=5, device=gpu)
    jobs = _jobs()
    stats = srv.run(jobs)
    mix = {j.jid: j for j in _finished(srv, stats)}
    for j in jobs:
        print(f"job {j.jid} {j.cls} n={j.n} steps={j.steps} admitted={j.admitted_at:.3f} done={j.done_at:.3f} "
              f"steps_done={j.steps_done} {j.result.get('error', '')}")
    print({k: v for k, v in stats.items() if not isinstance(v, list)})
    assert stats["jobs_finished"] == 8 and stats["jobs_rejected"] == 0
    waited = [j for j in srv.last_finished if j.admitted_at > 0]
    assert waited, "no job was admitted while others were running"
    assert 0 < stats["peak_resident_job_bytes"] <= stats["mem_budget_bytes"]
    for job in _jobs():
        alone_srv = Server(classes, capacity=CAP, max_systems=4, chunk=7, device=gpu)
        alone_srv.run([job])
        a = job.result
        b = mix[job.jid].result
        assert a["thermo"]["steps"] == job.steps == b["thermo"]["steps"]
        for key in "xvqp":
            assert np.array_equal(a[key], b[key]), (job.jid, key, float(np.abs(a[key] - b[key]).max()))
    # the plain path (single-system arena, device init) reaches the same states
    jobs = _jobs()
    st = run_plain(jobs, classes, chunk=6, device=gpu)
    assert st["jobs_finished"] == 8
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


def test_admission_waits_for_memory_budget(gpu):
    """A budget that holds the device bytes of one job but not two: the second job is admitted only
    after the first retires, and both finish."""
    from nepx.configs import make_system
    from nepx.serve import Job, Server
    classes = _classes(gpu)
    probe = Server(classes, capacity=2048, max_systems=4, chunk=5, device=gpu)
    a = Job(0, "tet10", make_system("tet10", 500, 0.2, seed=21), 12, 0.0)
    b = Job(1, "tet10", make_system("tet10", 500, 0.2, seed=22), 12, 0.0)
    one = probe.classes["tet10"].arena.job_bytes(a.spec)
    srv = Server(classes, capacity=2048, max_systems=4, chunk=5, device=gpu, mem_budget=int(1.5 * one))
    st = srv.run([a, b])
    assert st["jobs_finished"] == 2 and st["peak_resident_job_bytes"] <= int(1.5 * one)
    assert b.admitted_at > 0 and b.admitted_at >= a.done_at


def test_non_finite_job_and_cancelled_job_retire_alone(gpu):
    """Mid-run, one job's velocities are poisoned with a NaN and another job is cancelled: the first
    is stopped by the device at that step and retires with a fault, the second retires as cancelled,
    and the remaining jobs finish bitwise where they finish when served alone."""
    from nepx.configs import make_system
    from nepx.serve import Job, Server
    classes = _classes(gpu)

    def jobs():
        return [Job(0, "tet10", make_system("tet10", 300, 0.1, kT=0.8, seed=31), 25, 0.0),
                Job(1, "tet10", make_system("tet10", 400, 0.2, kT=1.2, seed=32), 25, 0.0),
                Job(2, "cube8", make_system("cube8", 216, 0.25, kT=1.0, seed=33, non_overlap=False), 25, 0.0),
                Job(3, "tet10", make_system("tet10", 350, 0.15, kT=1.0, seed=34), 400, 0.0)]
    js = jobs()
    calls = []

    def hook(server, now):
        calls.append(now)
        if len(calls) == 1:  # after the first chunk (5 steps)
            ar = server.classes["tet10"].arena
            a = int(ar.records()[0][js[1].slot]["p_start"])
            ar.t["vx"][a + 7] = float("nan")
            torch.cuda.synchronize()
            server.cancel(3)

    srv = Server(classes, capacity=2048, max_systems=4, chunk=5, device=gpu)
    st = srv.run(js, on_chunk=hook)
    for j in js:
        print(f"job {j.jid} steps_done={j.steps_done} {j.result.get('error', '')}")
    assert st["jobs_finished"] == 2 and st["jobs_failed"] == 2
    assert js[1].result.get("fault") and js[1].steps_done == 5
    assert "cancelled" in js[3].result["error"] and 5 <= js[3].steps_done < 400
    for j in (js[0], js[2]):
        ref = jobs()[j.jid]
        Server(classes, capacity=2048, max_systems=4, chunk=5, device=gpu).run([ref])
        for key in "xvqp":
            assert np.array_equal(ref.result[key], j.result[key]), (j.jid, key)


def test_collapsed_job_is_stopped_and_the_rest_continue(gpu):
    """A job whose 2,100 bodies sit in one small ball (every particle has more neighbours in the
    list radius than the build sorts) is stopped with an error at its first list build; the steps
    the overflow skipped were skipped by every job alike, and the others finish bitwise where they
    finish alone."""
    from nepx.configs import make_system
    from nepx.serve import Job, Server
    classes = _classes(gpu)
    blob = make_system("tet10", 2100, 0.05, seed=35)
    blob.x = np.random.default_rng(3).uniform(-1.0, 1.0, (2100, 3))

    def jobs():
        return [Job(0, "tet10", make_system("tet10", 300, 0.1, kT=0.8, seed=31), 20, 0.0),
                Job(1, "tet10", blob, 20, 0.0),
                Job(2, "cube8", make_system("cube8", 216, 0.25, kT=1.0, seed=33, non_overlap=False), 20, 0.0)]
    js = jobs()
    srv = Server(classes, capacity=4096, max_systems=4, chunk=5, device=gpu)
    st = srv.run(js)
    for j in js:
        print(f"job {j.jid} steps_done={j.steps_done} {j.result.get('error', '')}")
    assert st["jobs_finished"] == 2 and st["jobs_failed"] == 1
    assert "collapsed" in js[1].result["error"] and js[1].steps_done == 0
    for j in (js[0], js[2]):
        ref = jobs()[j.jid]
        Server(classes, capacity=4096, max_systems=4, chunk=5, device=gpu).run([ref])
        for key in "xvqp":
            assert np.array_equal(ref.result[key], j.result[key]), (j.jid, key)
