"""Batched ensemble advanced by graph replay (GPU): every replayed box must
equal its own eager single-box run bit for bit, including with eager work
and synchronisations between replays."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _specs():
    from fastpm.ensemble import config_specs

    # force meshes 16, 32, 48: all on the tiled (deterministic) paint
    return config_specs(n=5, seed=3, grids=(8, 16, 24), box_range=(50.0, 120.0))


def test_ensemble_graph_matches_eager(gpu):
    from fastpm.ensemble import Ensemble
    from fastpm.core import leapfrog

    stages = np.linspace(0.1, 1.0, 6)
    ens = Ensemble(_specs(), stages)
    ens.initialize()
    ens.run(use_graph=True)
    for b in ens.boxes:
        ref = b.solver.lpt(b.solver.linear(b.solver.whitenoise(b.seed), _pk(b)),
                           b.pm.generate_uniform_particle_grid(), stages[0])
        b.solver.nbody(ref, leapfrog(stages))
        for key in "SPF":
            assert torch.equal(getattr(b.state, key), getattr(ref, key)), (b.nc, key)
        assert b.state.a == ref.a


_PK = {}


def _pk(b):
    from fastpm.cosmology import LinearPower

    key = b.cosmo.sigma8
    if key not in _PK:
        _PK[key] = LinearPower(b.cosmo)
    return _PK[key]


def test_ensemble_replays_with_eager_work_between(gpu):
    """Replays interleaved with unrelated eager kernels and synchronisations
    give the same result as an eager ensemble."""
    from fastpm.ensemble import Ensemble

    stages = np.linspace(0.1, 1.0, 5)
    a = Ensemble(_specs(), stages)
    a.initialize()
    a.step_eager()
    torch.cuda.synchronize()
    a.capture()
    junk = torch.randn(1 << 20, device="cuda")
    for _ in range(len(stages) - 2):
        a.graph.replay()
        junk.mul_(1.0001).add_(1.0)
        torch.cuda.synchronize()
    a.finish()  # the last closing kick
    b = Ensemble(_specs(), stages)
    b.initialize()
    b.run(use_graph=False)
    for x, y in zip(a.boxes, b.boxes):
        for key in "SPF":
            assert torch.equal(getattr(x.state, key), getattr(y.state, key))


@pytest.mark.parametrize("resampler", ["cic", "tsc"])
def test_batched_step_equals_per_box_streams(gpu, resampler):
    """The batched step (fpm_batch_step: one launch per compiled kernel over
    all boxes) gives every box the same S, P, F and RHO, bit for bit, as the
    same ensemble run box by box on streams, eagerly and by graph replay."""
    from fastpm.ensemble import Ensemble, config_specs

    print("batched ensemble", resampler, flush=True)
    specs = config_specs(n=6, seed=5, grids=(16, 24, 32), box_range=(60.0, 200.0))
    stages = np.linspace(0.1, 1.0, 5)
    runs = []
    for batched, graph in ((True, True), (True, False), (False, False)):
        e = Ensemble(specs, stages, resampler=resampler, batched=batched)
        e.initialize()
        assert (e.plan is not None) == batched
        e.run(use_graph=graph)
        torch.cuda.synchronize()
        runs.append(e)
    ref = runs[-1]
    for e in runs[:-1]:
        for x, y in zip(e.boxes, ref.boxes):
            for key in ("S", "P", "F", "RHO"):
                assert torch.equal(getattr(x.state, key), getattr(y.state, key)), (x.nc, key)


def test_batch_plan_rejects_bad_boxes(gpu):
    import ctypes
    from fastpm import _lib

    L = _lib.lib()
    boxes = (_lib.FpmBatchBox * 1)()
    boxes[0].order = 4
    assert L.fpm_batch_device_bytes(1, boxes) < 0
    st = ctypes.c_int(0)
    mem = torch.empty(1 << 16, dtype=torch.uint8, device=gpu)
    assert not L.fpm_batch_create(1, boxes, mem.data_ptr(), mem.numel(), ctypes.byref(st))
    assert st.value < 0
    assert L.fpm_batch_device_bytes(0, boxes) < 0
