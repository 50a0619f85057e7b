"""muto_map_value_grad_noalloc (the MAP objective and its gradient in one
call that allocates nothing, FP64) against muto_map_forward +
muto_map_backward, through the C ABI: from stored lists Phi bit for bit;
with the paths walked again, where muto_map_forward sums each track's
moments per tile run and the no-alloc call in one lane per track, Phi within
1e-13 relative; the gradient up to the order of its FP64 additions; on two
calls in a row with the second on a smaller problem (fewer tracks and
voxels, so stale partial sums of the first call would show); sizes above
the static scratch are refused."""
import numpy as np
import pytest

import muontomography_amd as M
from muontomography_amd import workloads as W
from muontomography_amd._abi import make_grid, mlsd_cfg
from muontomography_amd.devmem import DeviceArray

pytestmark = pytest.mark.gpu


def _cfg(g):
    return mlsd_cfg(g, error_xy=0.5, error_dz_outer=500.0, error_dz_inner=500.0, update_with_mean=1,
                    mean_percentile=100, beta=0.7, p=1.3, delta=0.1, all_neighbours=1, add_object_prior=1)


def _lam(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.5e-6, 3e-6, n)


@pytest.mark.parametrize("recompute", [False, True])
def test_noalloc_matches_forward_backward(ctx, recompute):
    cases = [W.muons("small", n_muons=20_000), W.muons("fd32", n_muons=3_000)]
    assert cases[1][0].n_vox < cases[0][0].n_vox
    for k, (g, rays, _) in enumerate(cases):
        print(f"noalloc case {k}: {g.nx}^3 voxels, {rays.shape[0]} muons, recompute={recompute}", flush=True)
        plan = ctx.em_plan(_cfg(g), rays=rays)
        assert plan.counters()["n_scat_in_grid"] > 0
        lam = _lam(g.n_vox, 3 + k)
        sv = plan.map_forward(lam, M.MUTO_FP64, recompute=recompute)
        gref = sv.backward(1.0)
        sv.close()
        dl = DeviceArray.from_numpy(lam)
        dg = DeviceArray.from_numpy(np.full(g.n_vox, 5.0))  # overwritten, not added to
        dp = DeviceArray.from_numpy(np.array([np.nan]))
        plan.map_value_grad_noalloc(dl, dg, dp, recompute=recompute)
        phi = dp.numpy()[0]
        if recompute:
            assert abs(phi - sv.phi) <= 1e-13 * abs(sv.phi), (phi, sv.phi)
        else:
            assert phi == sv.phi, (phi, sv.phi)
        scale = np.abs(gref).max()
        assert scale > 0
        np.testing.assert_allclose(dg.numpy(), gref, rtol=0, atol=1e-13 * scale)
        # again on the same buffers: the same result (the scratch keeps nothing that matters)
        plan.map_value_grad_noalloc(dl, dg, dp, recompute=recompute)
        assert dp.numpy()[0] == phi
        np.testing.assert_allclose(dg.numpy(), gref, rtol=0, atol=1e-13 * scale)
        plan.close()


def test_noalloc_refuses_grids_beyond_its_scratch(ctx):
    g0, rays, _ = W.muons("small", n_muons=2_000)
    g = make_grid((257, 256, 256), g0.dx * g0.nx / 256, g0.x_min)  # 2^24 + 65536 voxels
    plan = ctx.em_plan(_cfg(g), rays=rays)
    dl = DeviceArray(g.n_vox)
    dg = DeviceArray(g.n_vox)
    dp = DeviceArray(1)
    with pytest.raises(M.MutoError, match="2\\^24"):
        plan.map_value_grad_noalloc(dl, dg, dp)
    plan.close()
