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
