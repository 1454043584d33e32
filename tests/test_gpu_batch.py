"""Batched replicas: several systems of different size, density and temperature advanced by the
same launches give bitwise the same states as each system run alone (no float atomics, block-aligned
per-system reductions, exact per-step pair sets)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
HYP = (11, 6, 4)


def _specs():
    from nepx.configs import make_system
    return [make_system("tet10", 300, 0.1, kT=0.7, seed=1),
            make_system("tet10", 777, 0.3, kT=1.6, seed=2),
            make_system("tet10", 1500, 0.45, kT=1.0, seed=3)]


def test_batched_equals_alone(gpu):
    import nepx
    from nepx.engine import StepRunner
    net = nepx.EnergyNet.random(40, [64, 64], seed=9, device=gpu)
    shape = nepx.get_shape("tet10")
    batched = nepx.Arena(_specs(), shape, HYP, net, device=gpu)
    rb = StepRunner(batched, use_graph=True)
    rb.init()
    rb.run(25)
    torch.cuda.synchronize()
    batched.check()
    for k, spec in enumerate(_specs()):
        alone = nepx.Arena([spec], shape, HYP, net, device=gpu)
        ra = StepRunner(alone, use_graph=False)
        ra.init()
        ra.run(25)
        torch.cuda.synchronize()
        a, b = alone.system_arrays(0), batched.system_arrays(k)
        for key in ("x", "v", "q", "p", "F", "T"):
            assert np.array_equal(a[key], b[key]), (k, key, float(np.abs(a[key] - b[key]).max()))
        ta, tb = alone.thermo()[0], batched.thermo()[k]
        assert ta["xi"] == tb["xi"] and ta["Kt"] == tb["Kt"] and ta["U"] == tb["U"]


def test_cube_replica_arena_equals_alone(gpu):
    """The batched-replica workload: 256 cube systems, N log-uniform in [64, 4096], packing fraction
    uniform in [0.2, 0.5], kT uniform in [0.5, 2.0] (seeded), stepped together with captured graphs;
    the smallest, the largest, the densest and two seeded others equal their runs alone bitwise."""
    import nepx
    from nepx.configs import replica_specs
    from nepx.engine import StepRunner
    specs = replica_specs(256, seed=2024)
    n = np.array([len(s.x) for s in specs])
    dens = np.array([len(s.x) / s.L ** 3 for s in specs])
    assert n.min() < 100 and n.max() > 3000 and n.sum() > 150_000
    pick = sorted({int(n.argmin()), int(n.argmax()), int(dens.argmax())} |
                  set(np.random.default_rng(7).choice(256, 2, replace=False).tolist()))
    net = nepx.EnergyNet.random(40, [64, 64], seed=13, device=gpu)
    shape = nepx.get_shape("cube8")
    steps = 20
    batched = nepx.Arena(specs, shape, HYP, net, device=gpu)
    rb = StepRunner(batched, use_graph=True)
    rb.init()
    rb.run(steps)
    torch.cuda.synchronize()
    batched.check()
    assert batched.device_steps() == steps
    print(f"cube replicas: {len(specs)} systems, {n.sum()} particles; checking systems {pick}")
    for k in pick:
        alone = nepx.Arena([specs[k]], shape, HYP, net, device=gpu)
        ra = StepRunner(alone, use_graph=True)
        ra.init()
        ra.run(steps)
        torch.cuda.synchronize()
        a, b = alone.system_arrays(0), batched.system_arrays(k)
        for key in ("x", "v", "q", "p", "F", "T"):
            assert np.array_equal(a[key], b[key]), (k, key, float(np.abs(a[key] - b[key]).max()))
        ta, tb = alone.thermo()[0], batched.thermo()[k]
        assert ta["xi"] == tb["xi"] and ta["Kt"] == tb["Kt"] and ta["U"] == tb["U"]
