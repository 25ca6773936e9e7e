"""fpm_power_bin_noalloc (block partials in the library's fixed device array)
against fpm_power_bin (caller workspace) through the C ABI (GPU only): equal
bit for bit on two calls in a row, the second on a smaller mesh with fewer
bins, for float32 and float64 fields, auto- and cross-spectra; a request
beyond the fixed array is refused before anything is launched."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _field(nmesh, dtype, seed, gpu):
    from fastpm.pm import ParticleMesh

    pm = ParticleMesh(Nmesh=nmesh, BoxSize=(300.0, 250.0, 200.0), dtype=dtype)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(nmesh).astype(dtype)
    return pm, pm.create("real", value=torch.from_numpy(x).to(gpu)).r2c()


def _both(pm, f1, f2, nbins, dtype):
    from fastpm import _lib
    from fastpm.pm import _stream

    L = _lib.lib()
    suf = _lib.suffix(dtype)
    kmax = np.pi * min(pm.Nmesh / pm.BoxSize)
    dk = kmax / nbins
    n = int(L.fpm_power_work_len(pm._desc, nbins))
    work = torch.empty(n, dtype=torch.float64, device=f1._t.device)
    ref = torch.full((nbins, 5), -1.0, dtype=torch.float64, device=f1._t.device)
    got = torch.full((nbins, 5), -2.0, dtype=torch.float64, device=f1._t.device)
    p2 = f2._t.data_ptr() if f2 is not None else None
    rc = getattr(L, "fpm_power_bin" + suf)(f1._t.data_ptr(), p2, pm._desc, pm.pitch2, nbins,
                                           0.0, dk, work.data_ptr(), n, ref.data_ptr(), _stream())
    assert rc == 0
    rc = getattr(L, "fpm_power_bin_noalloc" + suf)(f1._t.data_ptr(), p2, pm._desc, pm.pitch2,
                                                   nbins, 0.0, dk, got.data_ptr(), _stream())
    assert rc == 0
    torch.cuda.synchronize()
    return ref, got


@pytest.mark.parametrize("dtype", ["f4", "f8"])
def test_power_bin_noalloc_matches_workspace_call(gpu, dtype):
    print("noalloc power", dtype, flush=True)
    pm, a = _field((128, 96, 64), dtype, 1, gpu)
    _, b = _field((128, 96, 64), dtype, 2, gpu)
    for f2 in (None, b):
        ref, got = _both(pm, a, f2, 40, dtype)
        assert torch.equal(ref, got)
        assert float(ref[:, 0].sum()) > 0
    # second call: smaller mesh, fewer bins; the array keeps the first call's
    # partials beyond what this call writes, which it must not read
    pm2, c = _field((32, 16, 24), dtype, 3, gpu)
    ref, got = _both(pm2, c, None, 7, dtype)
    assert torch.equal(ref, got)


def test_power_bin_noalloc_rejects_oversize(gpu):
    from fastpm import _lib
    from fastpm.pm import _stream

    L = _lib.lib()
    m = _lib.FpmMesh()
    for d in range(3):
        m.nmesh[d] = 512
        m.box[d] = 1000.0
    pitch2 = 272
    dummy = torch.zeros(8, dtype=torch.complex64, device=gpu)
    out = torch.zeros((200, 5), dtype=torch.float64, device=gpu)
    # 2048 partial blocks at 512^3: 128 bins fit, 129 do not
    assert int(L.fpm_power_work_len(m, 129)) > 2048 * 128 * 5
    rc = L.fpm_power_bin_noalloc_f32(dummy.data_ptr(), None, m, pitch2, 129, 0.0, 0.01,
                                     out.data_ptr(), _stream())
    assert rc == -5
    assert b"scratch" in L.fpm_strerror(rc)
    torch.cuda.synchronize()
    assert torch.count_nonzero(out).item() == 0
