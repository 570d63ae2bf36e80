"""GPU tests of the batched solve of uneven meshes captured in one HIP graph:
every replay (with eager work and synchronisations between replays) equals
the eager run bit for bit and the FP64 solution of each mesh to 1e-6."""
import numpy as np
import pytest

from fluxpy_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def batch(torch_dev):
    from fluxpy_amd.batch import BatchedRadiosity
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    inst = meshes.batch_instances(count=4, nmin=400, nmax=3000, seed=77)
    models = [CgalTrimeshShapeModel(V.astype(np.float32), F) for V, F in inst]
    rng = np.random.default_rng(78)
    E_list, rho_list = [], []
    for i, sm in enumerate(models):
        D = meshes.sun_directions(4, elev_deg=(5.0, 30.0), seed=79 + i)
        E_list.append(sm.get_direct_irradiance_multi(1365.0, D).astype(np.float64))
        rho_list.append(rng.uniform(0.08, 0.3, sm.num_faces))
    br = BatchedRadiosity(models, E_list, rho_list, tol=1e-6, min_size=1024,
                          inner_iters=2, autotune=True)
    return br, models, E_list, rho_list


def test_batch_eager_matches_fp64(torch_dev, batch):
    from fluxpy_amd.solve import solve_radiosity
    br, models, E_list, rho_list = batch
    assert all(tn is not None for tn in br.tuning())
    outs = [B.clone() for B in br.run_eager()]
    ok, rel = br.verify(1e-7)
    assert ok, rel
    for it, B, E, rho in zip(br.items, outs, E_list, rho_list):
        Bref, _ = solve_radiosity(it["H"], E, rho)
        B = B.cpu().numpy()
        err = abs(B - Bref).max(0) / abs(Bref).max(0)
        assert err.max() <= 1e-6, err


def test_batch_graph_replays_match_eager(torch_dev, batch):
    t = torch_dev
    br, *_ = batch
    eager = [B.clone() for B in br.run_eager()]
    t.cuda.synchronize()
    br.capture()
    for rep in range(3):
        outs = br.replay()
        t.cuda.synchronize()
        for B, Be in zip(outs, eager):
            assert t.equal(B, Be), rep
        ok, rel = br.verify(1e-7)
        assert ok, (rep, rel)
        # eager work on the same buffers between replays
        again = br.run_eager()
        t.cuda.synchronize()
        for B, Be in zip(again, eager):
            assert t.equal(B, Be)
        junk = t.randn(1 << 20, device="cuda")
        junk.mul_(2.0)
        t.cuda.synchronize()
