"""fastpm.force.gravity.longrange / longrange_batch (reference
force/gravity.py:5-29) against the oracle's restatement (GPU only): split 0
and a Gaussian split, window compensation off and CIC, a factor != 1 and two
position sets, in float64 and float32."""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("split,deconv", [(0.0, 0), (4.0, 0), (2.5, 2)])
def test_longrange_batch_vs_oracle(gpu, dtype, split, deconv):
    from fastpm.pm import ParticleMesh, Particles
    from fastpm.force.gravity import longrange, longrange_batch

    print("longrange", dtype, split, deconv, flush=True)
    nmesh, box, factor = (32, 24, 16), (96.0, 72.0, 48.0), 0.45
    rng = np.random.default_rng(31)
    rho = 1 + 0.3 * rng.standard_normal(nmesh)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype)
    dk = pm.create("real", value=torch.from_numpy(rho.astype(dtype)).to(gpu)).r2c()
    # two position sets: explicit positions (some outside the box, wrapped)
    # and a lattice of the particle mesh with a displacement
    x1 = rng.uniform(-20, 120, size=(5000, 3)) * np.array(box) / 100.0
    pm_p = ParticleMesh(Nmesh=[n // 2 for n in nmesh], BoxSize=box, dtype=dtype)
    Q = pm_p.generate_uniform_particle_grid(shift=0.5)
    s2 = rng.normal(0, 2.0, size=(len(Q), 3))
    p1 = torch.from_numpy(x1.astype(dtype)).to(gpu)
    p2 = Particles(Q, torch.from_numpy(s2.astype(dtype)).to(gpu))
    got = longrange_batch([p1, p2], dk, split, factor, deconv=deconv)
    one = longrange(p1, dk, split, factor, deconv=deconv)
    # the oracle reads out at the positions the device holds (float32 rounding
    # of the inputs is not the force's error)
    X1 = p1.cpu().numpy().astype("f8")
    X2 = pm.positions(p2).cpu().numpy().astype("f8")
    ref = O.longrange([X1, X2], dk.value.cpu().numpy().astype("c16"), nmesh, box, split,
                      factor, 2, deconv)
    tol = 1e-10 if dtype == "f8" else 2e-5
    for g, r in zip(got, ref):
        assert g.shape == (len(r), 3) and g.dtype == pm.tdtype
        assert np.abs(g.cpu().numpy() - r).max() < tol * np.abs(r).max()
    assert torch.equal(one, got[0])
