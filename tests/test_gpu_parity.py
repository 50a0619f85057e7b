"""GPU parity: HIP kernels through the C ABI against the CPU FP64 oracle.

Contract: voxel index lists and PoCA assignments bit-exact; lengths and
PoCA points bit-exact (the contract allows 1e-12 relative); lambda images
bit-exact where the reference's order is deterministic (median, trimmed
mean on identical inputs), else within the tolerance stated per test.
"""
import os

import numpy as np
import pytest

import muontomography_amd as M
import oracle
from muontomography_amd import workloads as W
from muontomography_amd._abi import make_grid, mlsd_cfg, poca_cfg

pytestmark = pytest.mark.gpu


def _segments_siddon_test(n=10000):
    """The segments of the reference's test/SiddonTest.cpp:8-34 (libstdc++
    default_random_engine{0}, uniform over [-500, 500)), restated exactly."""
    import siddon_inputs
    return siddon_inputs.siddon_test_segments(n)


def _adversarial_segments(g):
    """Axis-parallel rays, rays through voxel corners/edges, rays starting or
    ending inside the grid, outside-parallel rays, degenerate segments."""
    lo = np.array([g.x_min, g.y_min, g.z_min])
    d = np.array([g.dx, g.dy, g.dz])
    n = np.array([g.nx, g.ny, g.nz], dtype=float)
    hi = lo + d * n
    P1, P2 = [], []
    rng = np.random.default_rng(7)
    for _ in range(300):  # plane-aligned endpoints (exact ties)
        i1 = rng.integers(-2, n.astype(int) + 3)
        i2 = rng.integers(-2, n.astype(int) + 3)
        P1.append(lo + d * i1)
        P2.append(lo + d * i2)
    for ax in range(3):  # axis-parallel, inside and outside
        for _ in range(50):
            a = lo + (hi - lo) * rng.uniform(-0.2, 1.2, 3)
            b = a.copy()
            b[ax] = lo[ax] + (hi[ax] - lo[ax]) * rng.uniform(-0.5, 1.5)
            P1.append(a)
            P2.append(b)
    for _ in range(300):  # start or end inside
        a = lo + (hi - lo) * rng.uniform(0, 1, 3)
        b = lo + (hi - lo) * rng.uniform(-1, 2, 3)
        if rng.uniform() < 0.5:
            a, b = b, a
        P1.append(a)
        P2.append(b)
    for _ in range(50):  # tiny and zero-length segments
        a = lo + (hi - lo) * rng.uniform(0, 1, 3)
        P1.append(a)
        P2.append(a + rng.normal(0, 1e-7, 3))
    P1.append(lo.copy()); P2.append(hi.copy())          # main diagonal
    P1.append(hi.copy()); P2.append(lo.copy())
    P1.append(lo + 0.5 * d); P2.append(lo + 0.5 * d)    # point
    return np.array(P1), np.array(P2)


def _check_siddon(ctx, g, p1, p2):
    ro, xo, yo, zo, lo_ = oracle.siddon_csr(g, p1, p2)
    rg, xg, yg, zg, lg = ctx.siddon_csr(g, p1, p2)
    assert np.array_equal(ro, rg), "per-segment voxel counts differ"
    assert np.array_equal(xo, xg) and np.array_equal(yo, yg) and np.array_equal(zo, zg)
    # lengths: bit-exact (FP64, no contraction, IEEE div/sqrt)
    assert np.array_equal(lo_.view(np.uint64), lg.view(np.uint64)), \
        f"max rel diff {np.max(np.abs(lo_ - lg) / np.maximum(np.abs(lo_), 1e-300))}"
    return int(ro[-1])


@pytest.mark.parametrize("grid", [
    make_grid(100, 2.0, -100.0),                 # test/SiddonTest.cpp:11-14
    make_grid((40, 30, 20), (25.0, 10.0, 7.5), (-500.0, -150.0, -80.0)),
    make_grid((7, 1, 3), (1.0, 3.0, 0.5), (0.0, 0.0, 0.0)),
    # nx above 2047: the fill stages x, y, z unpacked
    make_grid((2100, 3, 2), (0.1, 60.0, 90.0), (-105.0, -90.0, -90.0)),
])
def test_siddon_bitexact(ctx, grid):
    p1, p2 = _segments_siddon_test(5000)
    assert _check_siddon(ctx, grid, p1, p2) > 0 or grid.nx == 7
    a1, a2 = _adversarial_segments(grid)
    _check_siddon(ctx, grid, a1, a2)


def test_siddon_reference_test_golden(ctx):
    """GPU lists for SiddonTest.cpp's 10000 segments on its 100^3 x 2 mm grid
    equal the committed fixture (digests of counts, indices, length bits)."""
    import hashlib
    import json
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "siddon_test.json")) as f:
        fx = json.load(f)
    p1, p2 = _segments_siddon_test(fx["n_segments"])
    gd = fx["grid"]
    row, x, y, z, ln = ctx.siddon_csr(make_grid(gd["n"], gd["d"], gd["lo"]), p1, p2)
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    got = {"row_ptr_u64": h(np.asarray(row).astype("<u8")), "x_u32": h(np.asarray(x).astype("<u4")),
           "y_u32": h(np.asarray(y).astype("<u4")), "z_u32": h(np.asarray(z).astype("<u4")),
           "length_f64_bits": h(np.asarray(ln).astype("<f8"))}
    assert got == fx["digests"]


def test_poca_siddon_full_size_200(ctx):
    """Plain PoCA + Siddon at full size: 200^3 grid of 10 mm voxels, 1e7
    muons.  PoCA points, voxel casts and accept flags bit-exact for every
    muon; the Siddon lists of both legs (pin -> PoCA point -> pout) of every
    muon bit-exact by per-segment count and digest; full lists on a seeded
    1e5-muon subset."""
    g, rays, _ = W.muons("siddon200")
    assert len(rays) == 10_000_000
    Po, vo, ao, co = oracle.poca_points(g, rays)
    Pg, vg, ag, cg = ctx.poca_points(g, rays)
    assert np.array_equal(Po.view(np.uint64), Pg.view(np.uint64))
    assert np.array_equal(vo, vg) and np.array_equal(ao, ag)
    assert co["n_scat_in_grid"] == cg["n_scat_in_grid"] and co["n_not_parallel"] == cg["n_not_parallel"]
    total = 0
    for a, b in ((rays["pin"], Po), (Po, rays["pout"])):
        cg_, dg = ctx.siddon_digest(g, a, b)
        co_, do = oracle.siddon_digest(g, a, b)
        assert np.array_equal(cg_, co_), "per-segment counts differ"
        assert np.array_equal(dg, do), f"{int((dg != do).sum())} segment digests differ"
        total += int(co_.sum(dtype=np.uint64))
    assert total > 1_000_000_000
    sub = np.sort(np.random.default_rng(3).choice(len(rays), 100_000, replace=False))
    _check_siddon(ctx, g, rays["pin"][sub], Po[sub])
    _check_siddon(ctx, g, Po[sub], rays["pout"][sub])


def test_siddon_digest_matches_lists(ctx):
    """The digest entry point hashes exactly the lists muto_siddon_csr returns."""
    g = make_grid((40, 30, 20), (25.0, 10.0, 7.5), (-500.0, -150.0, -80.0))
    p1, p2 = _segments_siddon_test(3000)
    a1, a2 = _adversarial_segments(g)
    p1, p2 = np.concatenate([p1, a1]), np.concatenate([p2, a2])
    row, x, y, z, ln = ctx.siddon_csr(g, p1, p2)
    cnt, dig = ctx.siddon_digest(g, p1, p2)
    assert ctx.kernel_ms() > 0.0
    assert np.array_equal(cnt.astype(np.uint64), np.diff(row))
    mask = (1 << 64) - 1
    lb = ln.view(np.uint64)
    for i in range(0, len(p1), 37):
        h = 0xcbf29ce484222325
        for k in range(int(row[i]), int(row[i + 1])):
            for w in (int(x[k]) | (int(y[k]) << 32), int(z[k]), int(lb[k])):
                h = ((h ^ w) * 0x100000001b3) & mask
        assert h == int(dig[i])


def test_empty_and_fully_rejected_inputs(ctx, small_set):
    """No rays, and rays that all miss: PoCA, PoCut, SLR and the record
    reader return empty results as the oracle does; records that all hold a
    NaN give no rays."""
    g, rays, _ = small_set
    cfg = poca_cfg(g, length_weighted=1)
    none = rays[:0]
    for fn, ofn in ((ctx.poca_reconstruct, oracle.poca_reconstruct),
                    (lambda c, r: ctx.pocut_reconstruct(c, r, 0.5),
                     lambda c, r: oracle.pocut_reconstruct(c, r, 0.5)),
                    (ctx.slr_reconstruct, oracle.slr_reconstruct)):
        img, c = fn(cfg, none)
        assert not img.any() and c["n_total"] == 0
        # every track far away from the grid
        far = rays[:500].copy()
        for k in ("pin", "pout"):
            far[k][:, 0] += 1e6
        img, c = fn(cfg, far)
        io, co = ofn(cfg, far)
        assert not img.any() and not io.any()
        assert c["n_scat_in_grid"] == co["n_scat_in_grid"] == 0
    zl = [1500.0, 1000.0, 500.0, -500.0, -1000.0, -1500.0]
    rec = np.full((7, 14), np.nan)
    assert ctx.rays_from_records(rec, zl, 3).shape[0] == 0
    assert ctx.rays_from_records(rec[:0], zl, 3).shape[0] == 0
    with pytest.raises(M.MutoError):
        ctx.rays_from_records(np.zeros((3, 14)), zl, 1)  # one plane above


def test_siddon_empty_and_outside(ctx):
    g = make_grid(10, 1.0, 0.0)
    p1 = np.array([[20.0, 20.0, 20.0], [-5.0, 5.0, 5.0], [5.0, 5.0, 5.0]])
    p2 = np.array([[30.0, 25.0, 21.0], [-5.0, 5.0, 15.0], [5.0, 5.0, 5.0]])
    row, x, y, z, l = ctx.siddon_csr(g, p1, p2)
    ro, *_ = oracle.siddon_csr(g, p1, p2)
    assert np.array_equal(row, ro)
    assert row[1] == 0 and row[2] == 0
    r0, *_ = ctx.siddon_csr(g, np.zeros((0, 3)), np.zeros((0, 3)))
    assert r0.tolist() == [0]


def test_poca_points_bitexact(ctx, small_set):
    g, rays, _ = small_set
    for mode in (0, 1):
        Po, vo, ao, co = oracle.poca_points(g, rays, 1e-8, mode)
        Pg, vg, ag, cg = ctx.poca_points(g, rays, 1e-8, mode)
        assert np.array_equal(Po.view(np.uint64), Pg.view(np.uint64))
        assert np.array_equal(vo, vg)
        assert np.array_equal(ao, ag)
        assert co["n_not_parallel"] == cg["n_not_parallel"]
        assert co["n_scat_in_grid"] == cg["n_scat_in_grid"]


def test_poca_x86_cast_edges(ctx):
    """PoCA voxel casts below the grid wrap like x86 (B4): -0.5 -> 0,
    -1.5 -> 4294967295; both paths agree."""
    g = make_grid(10, 1.0, 0.0)
    rays = np.zeros(4, M.RAY_DTYPE)
    for i, x in enumerate([-0.5, -1.5, 9.99, 10.0]):
        rays[i]["pin"] = (x, 5.0, 9.0)
        rays[i]["pout"] = (x, 5.0, 1.0)
        rays[i]["din"] = (0.0, 0.1, 1.0)
        rays[i]["dout"] = (0.0, -0.1, 1.0)
    Po, vo, ao, _ = oracle.poca_points(g, rays)
    Pg, vg, ag, _ = ctx.poca_points(g, rays)
    assert np.array_equal(vo, vg) and np.array_equal(ao, ag)
    assert vo[0, 0] == 0 and vo[1, 0] == 4294967295


@pytest.mark.parametrize("lw,sl", [(1, 0), (0, 0), (1, 1), (0, 1)])
def test_poca_image(ctx, small_set, lw, sl):
    g, rays, _ = small_set
    cfg = poca_cfg(g, length_weighted=lw, straight_line=sl)
    io, co = oracle.poca_reconstruct(cfg, rays)
    ig, cg = ctx.poca_reconstruct(cfg, rays)
    assert co["n_scat_in_grid"] == cg["n_scat_in_grid"]
    assert np.array_equal(io == 0, ig == 0)
    # FP64 atomics reorder the per-voxel sums, and the scattering angle is a
    # difference of two atan2 values (ocml vs glibc differ by an ulp), which
    # cancellation amplifies for small angles: rtol 1e-9
    np.testing.assert_allclose(ig, io, rtol=1e-9, atol=0)


@pytest.mark.parametrize("lw,pct", [(1, 1.0), (1, 0.9), (0, 0.5), (1, 0.0), (0, 1.7)])
def test_pocut_image(ctx, small_set, lw, pct):
    """PoCut against the oracle (src/MutoPoCut.cpp:19-121): same tolerance
    as the PoCA image (atomic order of the weights, atan2 ulps)."""
    g, rays, _ = small_set
    cfg = poca_cfg(g, length_weighted=lw, straight_line=1)  # PoCut forces two legs
    io, co = oracle.pocut_reconstruct(cfg, rays, pct)
    ig, cg = ctx.pocut_reconstruct(cfg, rays, pct)
    for k in ("n_scat_in_grid", "n_not_parallel", "n_out_of_grid"):
        assert co[k] == cg[k], k
    assert np.array_equal(io == 0, ig == 0)
    np.testing.assert_allclose(ig, io, rtol=1e-9, atol=0)
    assert ctx.kernel_ms() > 0.0


@pytest.mark.parametrize("n", [4, 1])
def test_pocut_long_voxel_lists(ctx, small_set, n):
    """Coarse grids put many scattering points in a voxel: the lists go
    through the block sort in LDS (4^3: ~1e2 per voxel) and the global
    scratch sort (one voxel: ~8e3 points)."""
    g0, rays, _ = small_set
    g = make_grid(n, g0.dx * g0.nx / n, g0.x_min)
    cfg = poca_cfg(g, length_weighted=1)
    io, co = oracle.pocut_reconstruct(cfg, rays, 0.8)
    ig, cg = ctx.pocut_reconstruct(cfg, rays, 0.8)
    assert co["n_scat_in_grid"] == cg["n_scat_in_grid"]
    assert co["n_scat_in_grid"] / n ** 3 > (4096 if n == 1 else 16)
    np.testing.assert_allclose(ig, io, rtol=1e-9, atol=0)


@pytest.mark.parametrize("lw,he", [(1, 0), (0, 0), (1, 1)])
def test_slr_image(ctx, small_set, lw, he):
    """SLR against the oracle (src/MutoSLR.cpp:47-103)."""
    g, rays, _ = small_set
    cfg = poca_cfg(g, length_weighted=lw, has_energy=he)
    io, co = oracle.slr_reconstruct(cfg, rays)
    ig, cg = ctx.slr_reconstruct(cfg, rays)
    assert co["n_out_of_grid"] == cg["n_out_of_grid"] and co["n_total"] == cg["n_total"]
    assert np.array_equal(io == 0, ig == 0)
    np.testing.assert_allclose(ig, io, rtol=1e-9, atol=0)


def _em_cfg(g, **kw):
    base = dict(error_xy=0.5, error_dz_outer=500.0, error_dz_inner=500.0,
                initialize_method=2, initialize_value=1e-7, num_iteration=5)
    base.update(kw)
    return mlsd_cfg(g, **base)


def test_em_plan_paths_match_oracle(ctx, small_set):
    g, rays, _ = small_set
    cfg = _em_cfg(g)
    em = oracle.Em(cfg, rays)
    plan = ctx.em_plan(cfg, rays=rays)
    row, vox, ln, tl, t4, pr = plan.export()
    assert np.array_equal(row, em.row_ptr)
    assert np.array_equal(vox, em.vox)
    assert np.array_equal(ln.view(np.uint64), em.len.view(np.uint64))
    assert np.array_equal(tl.view(np.uint64), em.tail.view(np.uint64))
    # angles and displacements go through libm atan2/tan/cos: ulp-level only
    np.testing.assert_allclose(t4, em.trk4, rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(pr, em.p)
    mo = em.mj()
    assert np.array_equal(plan.voxel_counts(), mo)
    c = plan.counters()
    assert c["n_scat_in_grid"] == em.n_sel and c["nnz"] == em.nnz


@pytest.mark.parametrize("sl,he", [(1, 0), (0, 1), (1, 1)])
def test_mlsd_straight_line_and_energy(ctx, small_set, sl, he):
    """MutoMLSD's straight_line (the path is the one chord pin -> pout, not
    the two legs through the PoCA point: MutoMLSD.cpp:389) and has_energy
    (a momentum ratio per track, not one for all: MutoMLSD.cpp:189, 285,
    320) options: the plan's paths, tails and p_r match the oracle bit for
    bit (the angles and displacements, through libm, to 1e-13); on those
    prepared inputs four FP64 median iterations match it bit for bit; and
    10 FP32C mean-update iterations stay within 1e-4 of FP64."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=0, straight_line=sl, has_energy=he)
    em = oracle.Em(cfg, rays)
    plan = ctx.em_plan(cfg, rays=rays)
    row, vox, ln, tl, t4, pr = plan.export()
    assert np.array_equal(row, em.row_ptr) and np.array_equal(vox, em.vox)
    assert np.array_equal(ln.view(np.uint64), em.len.view(np.uint64))
    assert np.array_equal(tl.view(np.uint64), em.tail.view(np.uint64))
    np.testing.assert_allclose(t4, em.trk4, rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(pr, em.p)
    assert (len(np.unique(pr)) > 1) == bool(he)  # one ratio for all tracks without energies
    plan.close()
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    ref = img
    for _ in range(4):
        ref, _c, _ns = em.iterate(ref)
    plan.run(4)
    assert np.array_equal(plan.image().view(np.uint64), ref.view(np.uint64))
    plan.close()
    cfgm = _em_cfg(g, update_with_mean=1, mean_percentile=100, straight_line=sl, has_energy=he)
    p64 = ctx.em_plan(cfgm, rays=rays)
    p32 = ctx.em_plan(cfgm, rays=rays)
    p64.set_image(img)
    p32.set_image(img)
    c64 = p64.run(10, M.MUTO_FP64)
    c32 = p32.run(10, M.MUTO_FP32C)
    mx, rms, norm = _rel_stats(p32.image(), p64.image(), p64.voxel_counts())
    assert mx <= 1e-4, (mx, rms, norm)
    np.testing.assert_allclose(c32, c64, rtol=1e-5)
    p64.close()
    p32.close()


def test_mlsd_closed_form_tracks(ctx):
    """The constructed tracks of tests/mlsd_cases.py through the GPU plan:
    kinked tracks give the closed-form angles and displacements
    (src/MutoMLSD.cpp:350-376), PoCA points on the selection window's
    boundaries are accepted or rejected as src/MutoMLSD.cpp:298-301 says,
    and a kink inside one voxel, or an empty in-leg, gives the closed-form
    merged path (src/MutoMLSD.cpp:382-415).  The same inputs through the
    oracle match bit for bit (paths) or to 1e-13 (angles, through libm)."""
    import mlsd_cases as MC
    g = make_grid(20, 10.0, -100.0)
    for plane in (None, "xz", "yz"):
        rays, want = MC.random_kinks(g, 40, seed=11, in_plane=plane)
        cfg = _em_cfg(g)
        plan = ctx.em_plan(cfg, rays=rays)
        row, vox, ln, tl, t4, pr = plan.export()
        assert t4.shape == (40, 4)
        np.testing.assert_allclose(t4[:, :2], want[:, :2], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(t4[:, 2:], want[:, 2:], rtol=1e-9, atol=1e-9)
        em = oracle.Em(cfg, rays)
        np.testing.assert_allclose(t4, em.trk4, rtol=1e-13, atol=1e-15)
        assert np.array_equal(vox, em.vox) and np.array_equal(ln.view(np.uint64), em.len.view(np.uint64))
        plan.close()
    for gw, zdet in ((make_grid(10, 10.0, -50.0), (-1000.0, 1000.0)),
                     (make_grid(10, 100.0, -500.0), (-200.0, 200.0))):
        rays, want = MC.window_rays(gw, zdet)
        _, _, acc, _ = ctx.poca_points(gw, rays, mode=1)
        np.testing.assert_array_equal(acc, want)
        plan = ctx.em_plan(_em_cfg(gw), rays=rays)
        assert plan.counters()["n_scat_in_grid"] == int(want.sum())
        plan.close()
    for name, (rays, vox_w, ln_w, tail_w, _ni, _no) in MC.merge_cases().items():
        plan = ctx.em_plan(_em_cfg(MC.MERGE_GRID), rays=rays)
        row, vox, ln, tl, t4, pr = plan.export()
        np.testing.assert_array_equal(vox, vox_w)
        np.testing.assert_allclose(ln, ln_w, rtol=1e-12)
        np.testing.assert_allclose(tl, tail_w, rtol=1e-12, atol=1e-12)
        assert plan.counters()["n_empty_path"] == 0, name
        plan.close()


@pytest.mark.parametrize("mode", ["median", "trimmed", "mean"])
def test_em_fp64_prepared(ctx, small_set, mode):
    """Same prepared inputs: per-track Sigma, det, inverse and S_ij are
    bit-identical, so median and trimmed-mean images match bit for bit; the
    mean's per-voxel FP64 sum runs in atomic order (rel <= 1e-13)."""
    g, rays, _ = small_set
    kw = {"median": dict(update_with_mean=0), "trimmed": dict(update_with_mean=1, mean_percentile=98),
          "mean": dict(update_with_mean=1, mean_percentile=100)}[mode]
    cfg = _em_cfg(g, **kw)
    em = oracle.Em(cfg, rays)
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    ref = img
    costs = []
    for _ in range(4):
        ref, cost, _ns = em.iterate(ref)
        costs.append(cost)
    cg = plan.run(4)
    out = plan.image()
    np.testing.assert_allclose(cg, costs, rtol=1e-12)
    if mode == "mean":
        np.testing.assert_allclose(out, ref, rtol=1e-13, atol=0)
    else:
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), \
            f"max abs diff {np.max(np.abs(out - ref))}"


@pytest.mark.parametrize("mode", ["median", "trimmed"])
@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 2)])
def test_em_order_statistics_long_voxel_lists(ctx, small_set, mode, n):
    """Coarse grids over the same 1 m cube: 2^3 and 3x3x2 voxels of 333-500
    mm, so most voxels hold thousands of entries and several hold more than
    the 4096 that one block sorts in LDS (those go to the listed-voxel sort
    in global scratch, one block per voxel).  Median and trimmed mean
    (src/MutoMLSD.cpp:493-517) bit-exact against the oracle on its prepared
    inputs for 3 iterations (6e4 muons of the small set's generator)."""
    g0, rays, _ = W.muons("small", n_muons=60_000)
    ext = g0.dx * g0.nx
    g = make_grid(n, tuple(ext / k for k in n), g0.x_min)
    kw = dict(update_with_mean=0) if mode == "median" else dict(update_with_mean=1, mean_percentile=60)
    cfg = _em_cfg(g, **kw)
    em = oracle.Em(cfg, rays)
    mj = em.mj()
    assert (mj > 4096).sum() >= 2, mj
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    np.testing.assert_array_equal(plan.voxel_counts(), mj)
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    ref = img
    for _ in range(3):
        ref, _c, _ns = em.iterate(ref)
    plan.run(3)
    out = plan.image()
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), (out, ref, mj)


@pytest.mark.parametrize("mode", ["median", "mean"])
def test_em_singular_covariances(ctx, small_set, mode):
    """Covariances that are not invertible (MutoMLSD.cpp:456-463): with no
    measurement error and lambda = 0 over half the grid, the tracks that
    cross only that half have Sigma = 0.  Each such track counts a warning,
    adds no cost and leaves S_ij = 2 lambda_j (SigmaInv stays zero).  The
    GPU path matches the oracle's image and costs as the regular case does,
    and counts the same singular tracks."""
    g, rays, _ = small_set
    kw = dict(update_with_mean=0) if mode == "median" else dict(update_with_mean=1, mean_percentile=100)
    cfg = _em_cfg(g, error_xy=0.0, **kw)  # error matrix E = 0
    em = oracle.Em(cfg, rays)
    img = np.where(np.arange(g.n_vox) % g.nx < g.nx // 2, 1e-7, 0.0)
    ref, cost, ns = em.iterate(img)
    assert 0 < ns < em.n_sel, (ns, em.n_sel)
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    plan.set_image(img)
    cg = plan.run(1)
    out = plan.image()
    assert plan.counters()["n_singular"] == ns
    np.testing.assert_allclose(cg, [cost], rtol=1e-12)
    if mode == "mean":
        np.testing.assert_allclose(out, ref, rtol=1e-13, atol=0)
    else:
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    plan.close()


@pytest.mark.parametrize("mode", ["median", "trimmed", "mean"])
def test_em_fp64_wavefront_kernel(ctx, small_set, mode):
    """The wavefront-per-track FP64 kernel (MUTO_OPT_FP64_KERNEL 1) sums each
    covariance in a butterfly instead of the reference's running order, so
    it agrees with the oracle to FP64 rounding: images within 1e-10 relative
    after 4 iterations, costs within 1e-12."""
    g, rays, _ = small_set
    kw = {"median": dict(update_with_mean=0), "trimmed": dict(update_with_mean=1, mean_percentile=98),
          "mean": dict(update_with_mean=1, mean_percentile=100)}[mode]
    cfg = _em_cfg(g, **kw)
    em = oracle.Em(cfg, rays)
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    plan.set_option(M.MUTO_OPT_FP64_KERNEL, 1)
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    ref = img
    costs = []
    for _ in range(4):
        ref, cost, _ns = em.iterate(ref)
        costs.append(cost)
    cg = plan.run(4)
    out = plan.image()
    plan.close()
    np.testing.assert_allclose(cg, costs, rtol=1e-12)
    np.testing.assert_allclose(out, ref, rtol=1e-10, atol=0)
    assert np.array_equal(out == 0, ref == 0)


@pytest.mark.parametrize("mode", ["median", "trimmed", "mean"])
@pytest.mark.parametrize("brick", [(16, 16, 16, 256), (8, 16, 32, 512)])
def test_em_fp64_stream(ctx, small_set, mode, brick):
    """The FP64 brick stream (MUTO_OPT_FP64_KERNEL 2): the reference's exact
    lengths and tails in the brick-major stream, so every W_ij is the
    reference's and only the association of the covariance and mean sums
    differs.  Against the oracle on its prepared inputs, 4 iterations:
    images within 1e-10 relative (zeros where the oracle's are), costs within
    1e-12; median and trimmed mean sort each voxel's c_ij (one wavefront, or
    one block for longer lists)."""
    g, rays, _ = small_set
    kw = {"median": dict(update_with_mean=0), "trimmed": dict(update_with_mean=1, mean_percentile=60),
          "mean": dict(update_with_mean=1, mean_percentile=100)}[mode]
    cfg = _em_cfg(g, **kw)
    em = oracle.Em(cfg, rays)
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    plan.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
    plan.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    ref = img
    costs = []
    for _ in range(4):
        ref, cost, _ns = em.iterate(ref)
        costs.append(cost)
    cg = plan.run(4)
    out = plan.image()
    assert plan.counters()["n_segments"] > 0
    np.testing.assert_allclose(cg, costs, rtol=1e-12)
    np.testing.assert_allclose(out, ref, rtol=1e-10, atol=0)
    assert np.array_equal(out == 0, ref == 0)
    # graph replays and eager launches agree (the order statistics bit for bit)
    plan.set_image(img)
    plan.set_option(M.MUTO_OPT_GRAPH, 0)
    plan.run(4)
    eager = plan.image()
    if mode == "mean":  # FP64 LDS atomics: order-dependent at rounding level
        np.testing.assert_allclose(eager, out, rtol=1e-11, atol=0)
    else:
        assert np.array_equal(eager.view(np.uint64), out.view(np.uint64))
    plan.close()


@pytest.mark.parametrize("mode", ["median", "trimmed"])
def test_em_fp64_stream_long_voxel_lists(ctx, mode):
    """FP64 stream order statistics where voxels hold 2000-6800 entries (a
    2^3 grid over the 1 m cube, 6e4 muons): the one-block sorts in LDS and in
    the global scratch, within 1e-10 of the oracle after 3 iterations."""
    g0, rays, _ = W.muons("small", n_muons=60_000)
    ext = g0.dx * g0.nx
    g = make_grid(2, ext / 2, g0.x_min)
    kw = dict(update_with_mean=0) if mode == "median" else dict(update_with_mean=1, mean_percentile=98)
    cfg = _em_cfg(g, **kw)
    em = oracle.Em(cfg, rays)
    assert (em.mj() > 4096).sum() >= 2
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    plan.set_option(M.MUTO_OPT_BRICK, M.brick_code((2, 2, 2, 256)))
    plan.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    ref = img
    for _ in range(3):
        ref, _c, _ns = em.iterate(ref)
    plan.run(3)
    np.testing.assert_allclose(plan.image(), ref, rtol=1e-10, atol=0)
    plan.close()


@pytest.mark.parametrize("nb,p", [(0, 2.0), (1, 1.2)])
def test_map_fp64_stream(ctx, small_set, nb, p):
    """MAP iterations with the FP64 stream's ML step (mean update) against
    the reference-order FP64 plan: images within 1e-9, costs within 1e-12."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=p, delta=0.1,
                  all_neighbours=nb)
    img = np.full(g.n_vox, 1e-7)
    ref = ctx.em_plan(cfg, rays=rays)
    ref.set_image(img)
    c_ref = ref.run(6, M.MUTO_FP64, map_=True)
    plan = ctx.em_plan(cfg, rays=rays)
    plan.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
    plan.set_image(img)
    c = plan.run(6, M.MUTO_FP64, map_=True)
    np.testing.assert_allclose(c, c_ref, rtol=1e-12)
    np.testing.assert_allclose(plan.image(), ref.image(), rtol=1e-9, atol=0)
    ref.close()
    plan.close()


def test_mlsd_one_shot(ctx, small_set):
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=0, num_iteration=3)
    io, co, no = oracle.mlsd_reconstruct(cfg, rays)
    ig, cg, ng = ctx.mlsd_reconstruct(cfg, rays)
    assert no["n_scat_in_grid"] == ng["n_scat_in_grid"] and no["nnz"] == ng["nnz"]
    assert ng["iterations_done"] == 3
    # inputs differ from the oracle only by libm ulps in the angles
    np.testing.assert_allclose(ig, io, rtol=1e-9, atol=1e-22)
    np.testing.assert_allclose(cg, co, rtol=1e-9)


@pytest.mark.parametrize("prec", [M.MUTO_FP64, M.MUTO_FP64_STREAM])
@pytest.mark.parametrize("mean", [0, 1])
def test_mlsd_reference_config_poca_init_20_iterations(ctx, mean, prec):
    """The CPU reference configuration in full: 40^3 voxels of 25 mm, 1e5
    muons, the PoCA image (length-weighted, two legs) handed over in FP64 as
    the initial image, then 20 MLSD iterations, median (the reference's
    default) and mean, in the reference's order and on the FP64 brick stream
    (MUTO_FP64_STREAM, through the one-shot call).  Inputs differ from the
    oracle's only by libm ulps in the angles (and the stream's sums by
    association); after 20 iterations the images agree within 1e-8."""
    g, rays, _ = W.muons("small")
    assert len(rays) == 100_000
    pc = poca_cfg(g, length_weighted=1, straight_line=0)
    po, _ = oracle.poca_reconstruct(pc, rays)
    pg, _ = ctx.poca_reconstruct(pc, rays)
    np.testing.assert_allclose(pg, po, rtol=1e-9, atol=0)
    cfg = _em_cfg(g, update_with_mean=mean, mean_percentile=100, num_iteration=20, initialize_method=0)
    io, co, no = oracle.mlsd_reconstruct(cfg, rays, init=po)
    cfg.precision = prec
    ig, cg, ng = ctx.mlsd_reconstruct(cfg, rays, init=pg)
    assert ng["iterations_done"] == no["iterations_done"] == 20
    assert ng["n_scat_in_grid"] == no["n_scat_in_grid"] and ng["nnz"] == no["nnz"]
    np.testing.assert_allclose(ig, io, rtol=1e-8, atol=1e-22)
    np.testing.assert_allclose(cg, co, rtol=1e-9)


def _rel_stats(a, ref, mj):
    m = (mj > 0) & (ref > 0)
    r = np.abs(a[m] - ref[m]) / ref[m]
    return float(r.max()), float(np.sqrt(np.mean(r * r))), float(np.abs(a - ref).max() / np.abs(ref).max())


def test_fp32c_matches_fp64(ctx, small_set):
    """FP32C (FP32 entry terms, FP64 and int64 fixed-point sums, FP64 2x2
    algebra) stays within 1e-4 relative of FP64 after 20 iterations, mean
    update."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
    p64 = ctx.em_plan(cfg, rays=rays)
    p32 = ctx.em_plan(cfg, rays=rays)
    img = np.full(g.n_vox, 1e-7)
    p64.set_image(img)
    p32.set_image(img)
    c64 = p64.run(20, M.MUTO_FP64)
    c32 = p32.run(20, M.MUTO_FP32C)
    a64, a32 = p64.image(), p32.image()
    mx, rms, norm = _rel_stats(a32, a64, p64.voxel_counts())
    assert mx <= 1e-4, (mx, rms, norm)
    np.testing.assert_allclose(c32, c64, rtol=1e-5)
    assert p32.counters()["n_segments"] > 0


@pytest.mark.parametrize("mode", ["median", "trimmed"])
@pytest.mark.parametrize("brick", [(16, 16, 16, 256), (8, 16, 32, 512)])
def test_fp32c_order_statistics_match_fp64(ctx, small_set, mode, brick):
    """FP32C median and trimmed mean (src/MutoMLSD.cpp:493-517: each c_ij
    from the stream's FP32 lengths written to its voxel's slot as FP32, the
    order statistic selected among the voxel's values) against the
    reference-order FP64 run after 20 iterations: voxels crossed by 8 or more
    tracks within 1e-4 relative, every voxel within 5e-4 (measured: median
    1.9e-5, 60% trimmed mean 1.3e-4, both at voxels that one or two tracks
    cross, whose values grow 1e6-fold over the run), costs within 1e-5;
    graph replays equal eager launches and a second run bit for bit (every
    c_ij has its own slot, the selection does not depend on the write
    order)."""
    g, rays, _ = small_set
    kw = dict(update_with_mean=0) if mode == "median" else dict(update_with_mean=1, mean_percentile=60)
    cfg = _em_cfg(g, **kw)
    p64 = ctx.em_plan(cfg, rays=rays)
    p32 = ctx.em_plan(cfg, rays=rays)
    p32.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
    img = np.full(g.n_vox, 1e-7)
    p64.set_image(img)
    p32.set_image(img)
    c64 = p64.run(20, M.MUTO_FP64)
    c32 = p32.run(20, M.MUTO_FP32C)
    a64, a32 = p64.image(), p32.image()
    assert np.abs(a64).max() > 0
    mj = p64.voxel_counts()
    mx, rms, norm = _rel_stats(a32, a64, mj)
    assert mx <= 5e-4, (mx, rms, norm)
    mx8, _, _ = _rel_stats(a32, a64, np.where(mj >= 8, mj, 0))
    assert mx8 <= 1e-4, (mx8, mx, rms)
    assert np.array_equal(a32 == 0, a64 == 0)
    np.testing.assert_allclose(c32, c64, rtol=1e-5)
    p32.set_image(img)
    p32.run(20, M.MUTO_FP32C)
    assert np.array_equal(p32.image().view(np.uint64), a32.view(np.uint64))
    p32.set_image(img)
    p32.set_option(M.MUTO_OPT_GRAPH, 0)
    p32.run(20, M.MUTO_FP32C)
    assert np.array_equal(p32.image().view(np.uint64), a32.view(np.uint64))
    p64.close()
    p32.close()


@pytest.mark.parametrize("mode", ["median", "trimmed"])
def test_fp32c_order_statistics_long_voxel_lists(ctx, mode):
    """FP32C order statistics where voxels hold 2000-6800 entries (a 2^3 grid
    over the 1 m cube, 6e4 muons: the one-block sorts in LDS and in the
    global scratch) within 1e-4 of the reference-order FP64 run after 3
    iterations."""
    g0, rays, _ = W.muons("small", n_muons=60_000)
    ext = g0.dx * g0.nx
    g = make_grid(2, ext / 2, g0.x_min)
    kw = dict(update_with_mean=0) if mode == "median" else dict(update_with_mean=1, mean_percentile=98)
    cfg = _em_cfg(g, **kw)
    p64 = ctx.em_plan(cfg, rays=rays)
    p32 = ctx.em_plan(cfg, rays=rays)
    assert (p64.voxel_counts() > 4096).sum() >= 2
    p32.set_option(M.MUTO_OPT_BRICK, M.brick_code((2, 2, 2, 256)))
    img = np.full(g.n_vox, 1e-7)
    p64.set_image(img)
    p32.set_image(img)
    p64.run(3, M.MUTO_FP64)
    p32.run(3, M.MUTO_FP32C)
    mx, rms, norm = _rel_stats(p32.image(), p64.image(), p64.voxel_counts())
    assert mx <= 1e-4, (mx, rms, norm)
    if mode == "median":
        # brackets (the default) and whole lists in slots pick the same values
        a = p32.image()
        p32.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, 0)
        p32.set_image(img)
        p32.run(3, M.MUTO_FP32C)
        assert np.array_equal(p32.image().view(np.uint64), a.view(np.uint64))
    p64.close()
    p32.close()


@pytest.mark.parametrize("brick", [(16, 16, 16, 256), (8, 16, 32, 512), (8, 8, 16, 1024)])
def test_fp32c_median_brackets_equal_slots(ctx, small_set, brick):
    """FP32C median by brackets (MUTO_OPT_MEDIAN_BRACKETS 1: each voxel's
    median picked among the values inside a bracket around its last one, the
    voxels whose median left it through their whole lists) against every
    c_ij written to its slot (option 0): both pick the same FP32 values with
    the same arithmetic, so the images agree bit for bit after each of 12
    iterations (src/MutoMLSD.cpp:505-515), in a second run that starts from
    the first run's brackets, without the graph, and for MAP iterations."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=0, beta=1.0, p=1.2, delta=0.1, all_neighbours=1,
                  add_object_prior=0)
    pa = ctx.em_plan(cfg, rays=rays)
    pb = ctx.em_plan(cfg, rays=rays)
    for p in (pa, pb):
        p.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
    pb.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, 0)
    img = np.full(g.n_vox, 1e-7)
    pa.set_image(img)
    pb.set_image(img)
    for it in range(12):
        ca = pa.run(1, M.MUTO_FP32C)
        cb = pb.run(1, M.MUTO_FP32C)
        assert np.array_equal(pa.image().view(np.uint64), pb.image().view(np.uint64)), it
        np.testing.assert_allclose(ca, cb, rtol=1e-12)  # (per-block cost sums meet in any order)
    assert np.abs(pa.image()).max() > 0
    ref = pb.image()
    pb.set_image(img)
    pb.run(12, M.MUTO_FP32C)
    assert np.array_equal(pb.image().view(np.uint64), ref.view(np.uint64))
    for graph in (1, 0):
        pa.set_option(M.MUTO_OPT_GRAPH, graph)
        pa.set_image(img)
        pa.run(12, M.MUTO_FP32C)
        assert np.array_equal(pa.image().view(np.uint64), ref.view(np.uint64)), graph
    pa.set_image(img)
    pb.set_image(img)
    pa.run(4, M.MUTO_FP32C, map_=True)
    pb.run(4, M.MUTO_FP32C, map_=True)
    assert np.array_equal(pa.image().view(np.uint64), pb.image().view(np.uint64))
    pa.close()
    pb.close()


@pytest.mark.parametrize("brick", [(8, 16, 32, 512), (16, 16, 16, 256)])
def test_fp64_stream_median_brackets_equal_slots(ctx, small_set, brick):
    """The FP64 stream's median by brackets (the exact FP64 c_ij classified
    against FP32 bounds rounded outwards) against every value in its slot:
    the same images bit for bit after each of 8 iterations and for MAP
    iterations."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=0, beta=1.0, p=1.2, delta=0.1, all_neighbours=1,
                  add_object_prior=0)
    pa = ctx.em_plan(cfg, rays=rays)
    pb = ctx.em_plan(cfg, rays=rays)
    for p in (pa, pb):
        p.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
        p.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
    pb.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, 0)
    img = np.full(g.n_vox, 1e-7)
    pa.set_image(img)
    pb.set_image(img)
    for it in range(8):
        pa.run(1, M.MUTO_FP64)
        pb.run(1, M.MUTO_FP64)
        assert np.array_equal(pa.image().view(np.uint64), pb.image().view(np.uint64)), it
    assert np.abs(pa.image()).max() > 0
    pa.set_image(img)
    pb.set_image(img)
    pa.run(4, M.MUTO_FP64, map_=True)
    pb.run(4, M.MUTO_FP64, map_=True)
    assert np.array_equal(pa.image().view(np.uint64), pb.image().view(np.uint64))
    pa.close()
    pb.close()


@pytest.mark.parametrize("prec", ["fp32c", "fp64_stream"])
def test_median_voxels_beyond_bracket_counts(ctx, prec):
    """Voxels of more than 65535 entries (a 1x1x2 grid over the 1 m cube, 3e5
    muons: about 8.4e4 entries each), beyond the 16-bit counts of the
    brackets: the median runs on every value in its slot whatever
    MUTO_OPT_MEDIAN_BRACKETS says, so both settings give the same image, with
    no misses counted, within the precision's bound of the oracle after 3
    iterations (src/MutoMLSD.cpp:505-515)."""
    g0, rays, _ = W.muons("small", n_muons=300_000)
    ext = g0.dx * g0.nx
    g = make_grid((1, 1, 2), (ext, ext, ext / 2), g0.x_min)
    cfg = _em_cfg(g, update_with_mean=0)
    em = oracle.Em(cfg, rays)
    assert em.mj().min() >= 65536, em.mj()
    img = np.full(g.n_vox, 1e-7)
    ref = img
    for _ in range(3):
        ref, _c, _ns = em.iterate(ref)
    out = []
    for brk in (1, 0):
        plan = ctx.em_plan(cfg, rays=rays)
        plan.set_option(M.MUTO_OPT_BRICK, M.brick_code((2, 2, 2, 256)))
        plan.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, brk)
        if prec == "fp64_stream":
            plan.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
        plan.set_image(img)
        plan.run(3, M.MUTO_FP32C if prec == "fp32c" else M.MUTO_FP64)
        assert plan.counters()["n_median_miss"] == 0
        out.append(plan.image())
        plan.close()
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))
    np.testing.assert_allclose(out[0], ref, rtol=1e-4 if prec == "fp32c" else 1e-10, atol=0)


def test_map_fp32c_median_close_to_fp64(ctx, small_set):
    """MAP iterations whose ML step is the median (the reference's default
    update rule) in FP32C against FP64: images within 1e-4 relative after 8
    iterations."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=0, beta=1.0, p=1.2, delta=0.1, all_neighbours=1,
                  add_object_prior=0, num_iteration=8)
    p64 = ctx.em_plan(cfg, rays=rays)
    p32 = ctx.em_plan(cfg, rays=rays)
    img = np.full(g.n_vox, 1e-7)
    p64.set_image(img)
    p32.set_image(img)
    c64 = p64.run(8, M.MUTO_FP64, map_=True)
    c32 = p32.run(8, M.MUTO_FP32C, map_=True)
    a64, a32 = p64.image(), p32.image()
    assert np.all(np.isfinite(a32)) and np.abs(a64).max() > 0
    mx, rms, norm = _rel_stats(a32, a64, p64.voxel_counts())
    assert mx <= 1e-4, (mx, rms, norm)
    np.testing.assert_allclose(c32, c64, rtol=1e-5)
    p64.close()
    p32.close()


@pytest.mark.parametrize("shape", [(16, 16, 16, 256), (16, 16, 32, 512)])
def test_fp32c_two_sum_mode(ctx, small_set, shape):
    """MUTO_OPT_ACC_MODE 0 (FP32 two-sum sums in both entry passes, FP64
    atomics at the flush): within 1e-4 of FP64 after 20 mean-update
    iterations, as the default mode is, with 256- and 512-thread chunks (the
    backprojection then takes two chunks per workgroup)."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
    p64 = ctx.em_plan(cfg, rays=rays)
    p32 = ctx.em_plan(cfg, rays=rays)
    p32.set_option(M.MUTO_OPT_ACC_MODE, 0)
    p32.set_option(M.MUTO_OPT_BRICK, M.brick_code(shape))
    img = np.full(g.n_vox, 1e-7)
    p64.set_image(img)
    p32.set_image(img)
    c64 = p64.run(20, M.MUTO_FP64)
    c32 = p32.run(20, M.MUTO_FP32C)
    mx, rms, norm = _rel_stats(p32.image(), p64.image(), p64.voxel_counts())
    assert mx <= 1e-4, (mx, rms, norm)
    np.testing.assert_allclose(c32, c64, rtol=1e-5)
    p64.close()
    p32.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fp32c_one_voxel_axis(ctx, small_set, axis):
    """Grids one voxel wide along an axis (40 x 40 voxels of 25 mm across
    the others, one 1000 mm voxel along it), where the packed entries' step
    codes must tell the remaining axes apart (a step of +-nx is a y step
    when nx = 1): FP32C within 1e-4 of FP64 after 10 mean-update
    iterations, with the FP64 plan's merged paths checked against the
    oracle's first."""
    g0, rays, _ = small_set
    n = [40, 40, 40]
    d = [25.0, 25.0, 25.0]
    n[axis], d[axis] = 1, 1000.0
    g = make_grid(tuple(n), tuple(d), -500.0)
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
    em = oracle.Em(cfg, rays)
    assert em.n_sel > 1000
    p64 = ctx.em_plan(cfg, rays=rays)
    row, vox, ln, tail, trk4, pr = p64.export()
    np.testing.assert_array_equal(row, em.row_ptr)
    np.testing.assert_array_equal(vox, em.vox)
    p32 = ctx.em_plan(cfg, rays=rays)
    p32.set_option(M.MUTO_OPT_BRICK, M.brick_code((16, 16, 32, 512)))
    img = np.full(g.n_vox, 1e-7)
    p64.set_image(img)
    p32.set_image(img)
    c64 = p64.run(10, M.MUTO_FP64)
    c32 = p32.run(10, M.MUTO_FP32C)
    mx, rms, norm = _rel_stats(p32.image(), p64.image(), p64.voxel_counts())
    assert mx <= 1e-4, (mx, rms, norm)
    np.testing.assert_allclose(c32, c64, rtol=1e-5)
    assert p32.counters()["n_segments"] > 0
    p64.close()
    p32.close()


def test_fp32c_graph_replays_match_eager(ctx, small_set):
    """The captured FP32C iteration graph, replayed one iteration at a time
    with host reads in between and then several iterations back to back,
    gives the images of the eager launches bit for bit and their costs up
    to the order of the FP64 cost sum (rtol 1e-9)."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
    img = np.full(g.n_vox, 1e-7)
    out = {}
    for graph in (0, 1):
        plan = ctx.em_plan(cfg, rays=rays)
        plan.set_option(M.MUTO_OPT_GRAPH, graph)
        plan.set_image(img)
        steps = []
        for _ in range(3):
            c = plan.run(1, M.MUTO_FP32C)
            steps.append((plan.image(), c))
        c4 = plan.run(4, M.MUTO_FP32C)
        steps.append((plan.image(), c4))
        out[graph] = steps
        plan.close()
    for (a_e, c_e), (a_g, c_g) in zip(out[0], out[1]):
        # images: bit-identical (fixed-point sums); costs: FP64 atomic order
        assert np.array_equal(a_g.view(np.uint64), a_e.view(np.uint64))
        np.testing.assert_allclose(c_g, c_e, rtol=1e-9)


@pytest.mark.parametrize("brick", [(16, 16, 16, 256), (8, 8, 8, 256)])
def test_fp32c_image_is_deterministic(ctx, small_set, brick):
    """FP32C images are the same bits run to run, also where several tasks
    share a brick (their int64 sums are combined exactly by the last one)."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
    img = np.full(g.n_vox, 1e-7)
    outs = []
    for _ in range(2):
        plan = ctx.em_plan(cfg, rays=rays)
        plan.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
        plan.set_image(img)
        plan.run(6, M.MUTO_FP32C)
        outs.append(plan.image())
        plan.set_image(img)
        plan.run(6, M.MUTO_FP32C)
        outs.append(plan.image())
        plan.close()
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))


def test_fp32c_precision_em256(ctx):
    """FP32C against FP64 on the 256^3 x 10 mm grid with 2e6 muons (about
    3.6e8 entries, 2.3e7 brick segments): 40 mean-update iterations from a
    uniform start.

    With 2e6 muons some voxels hold about ten tracks, and there the iteration
    amplifies any perturbation: FP64 itself, started from the uniform image
    perturbed by a relative 1e-8 (1e-7), ends up to 2e-5 (3e-4) away from the
    unperturbed FP64 run after 40 iterations, in such a voxel.  The reference
    for the error is therefore FP64 EM whose image is rounded to FP32 before
    every iteration (the precision of an image held in FP32, as the forward
    gather reads it): FP32C's RMS error over voxels must not exceed twice
    that run's, nor its maximum four times that run's (or 1e-4 when that is
    larger); costs within 1e-6.  (The 1e-4 bound on the maximum itself holds
    at the 1e7-muon bench configuration, which bench.py reports.)"""
    from muontomography_amd import workloads as W
    g, rays, _ = W.muons("em256", n_muons=2_000_000)
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
    plan = ctx.em_plan(cfg, rays=rays)
    mj = plan.voxel_counts()
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    c64 = plan.run(40, M.MUTO_FP64)
    a64 = plan.image()
    lam = img.copy()
    for _ in range(40):
        plan.set_image(lam.astype(np.float32).astype(np.float64))
        plan.run(1, M.MUTO_FP64)
        lam = plan.image()
    f_mx, f_rms, _ = _rel_stats(lam, a64, mj)
    plan.set_image(img)
    c32 = plan.run(40, M.MUTO_FP32C)
    a32 = plan.image()
    mx, rms, norm = _rel_stats(a32, a64, mj)
    print(f"fp32c em256 2e6: max {mx:.3e} rms {rms:.3e} norm {norm:.3e}; "
          f"fp64 with an FP32 image: max {f_mx:.3e} rms {f_rms:.3e}", flush=True)
    assert rms <= 2.0 * f_rms, (mx, rms, f_mx, f_rms)
    assert mx <= max(1e-4, 4.0 * f_mx), (mx, rms, f_mx, f_rms)
    np.testing.assert_allclose(c32, c64, rtol=1e-6)
    plan.close()


@pytest.mark.parametrize("nb,obj,p", [(0, 0, 1.0), (1, 1, 2.0)])
def test_map_fp64(ctx, small_set, nb, obj, p):
    """MAP (MutoMAP.cpp): ML step + per-voxel cubic surrogate step.  On the
    same prepared inputs the ML part is bit-exact; the voxel step uses pow()
    and log(), which differ from glibc by ulps, so rtol 1e-9."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=0, beta=1.0, p=p, delta=0.1, all_neighbours=nb,
                  add_object_prior=obj, num_iteration=3)
    em = oracle.Em(cfg, rays)
    plan = ctx.em_plan(cfg, prepared=em.prepared())
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    mj = em.mj()
    ref = img.copy()
    for _ in range(3):
        ml, _c, _ = em.iterate(ref)
        ref, _w = oracle.map_voxel_step(cfg, mj, ref, ml)
    plan.run(3, M.MUTO_FP64, map_=True)
    out = plan.image()
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-22)
    # one-shot MAP from rays
    io, co, no = oracle.mlsd_reconstruct(cfg, rays, map_=True)
    ig, cg, ng = ctx.map_reconstruct(cfg, rays)
    np.testing.assert_allclose(ig, io, rtol=1e-8, atol=1e-20)
    assert ng["n_cubic_warn"] == no["n_cubic_warn"]


@pytest.mark.parametrize("nb,p", [(0, 2.0), (1, 1.2)])
def test_map_fp32c_iterations_close_to_fp64(ctx, small_set, nb, p):
    """MAP iterations (MutoMAP.cpp:39-205) in FP32C -- the mean ML step from
    the brick stream, then the FP64 voxel step on its image -- against the
    same iterations in FP64: images within 1e-4 relative after 8 iterations,
    costs (of the ML step) within 1e-5."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=p, delta=0.1,
                  all_neighbours=nb, add_object_prior=0, num_iteration=8)
    p64 = ctx.em_plan(cfg, rays=rays)
    p32 = ctx.em_plan(cfg, rays=rays)
    img = np.full(g.n_vox, 1e-7)
    p64.set_image(img)
    p32.set_image(img)
    c64 = p64.run(8, M.MUTO_FP64, map_=True)
    c32 = p32.run(8, M.MUTO_FP32C, map_=True)
    a64, a32 = p64.image(), p32.image()
    assert np.all(np.isfinite(a32)) and np.abs(a64).max() > 0
    mx, rms, norm = _rel_stats(a32, a64, p64.voxel_counts())
    assert mx <= 1e-4, (mx, rms, norm)
    np.testing.assert_allclose(c32, c64, rtol=1e-5)
    # the MAP step moved the image away from the plain ML result
    p_ml = ctx.em_plan(cfg, rays=rays)
    p_ml.set_image(img)
    p_ml.run(8, M.MUTO_FP32C)
    assert np.abs(p_ml.image() - a32).max() > 1e-3 * np.abs(a32).max()
    for q in (p64, p32, p_ml):
        q.close()


def test_early_stop_map_from_ml_image(ctx, small_set):
    """early_stop (MutoMLSD.cpp:209-212, MutoMAP.cpp:194-197): MAP started
    from a converged ML image raises the ML cost at its second iteration, so
    the run stops there and keeps that iteration's image, as the oracle
    does.  One-shot FP64 against the oracle; FP64 and FP32C plans stop at
    the same iteration."""
    g, rays, _ = small_set
    base = dict(update_with_mean=1, mean_percentile=100)
    plan = ctx.em_plan(_em_cfg(g, num_iteration=60, **base), rays=rays)
    plan.set_image(np.full(g.n_vox, 1e-7))
    plan.run(60, M.MUTO_FP64)
    ml = plan.image()
    plan.close()
    cfg = _em_cfg(g, num_iteration=10, beta=1.0, p=2.0, delta=0.1, early_stop=1,
                  initialize_method=0, **base)
    io, co, no = oracle.mlsd_reconstruct(cfg, rays, init=ml, map_=True)
    assert no["iterations_done"] == 2 and co[1] > co[0], co
    ig, cg, ng = ctx.map_reconstruct(cfg, rays, init=ml)
    assert ng["iterations_done"] == 2
    np.testing.assert_allclose(cg, co, rtol=1e-9)
    np.testing.assert_allclose(ig, io, rtol=1e-8, atol=1e-20)
    for prec, tol in ((M.MUTO_FP64, 1e-8), (M.MUTO_FP32C, 1e-4)):
        p = ctx.em_plan(cfg, rays=rays)
        p.set_image(ml)
        c = p.run(10, prec, map_=True)
        assert len(c) == 2, (prec, c)
        np.testing.assert_allclose(c, co, rtol=tol)
        mx, rms, norm = _rel_stats(p.image(), io, p.voxel_counts())
        assert mx <= tol, (prec, mx, rms, norm)
        p.close()


def test_initial_image_methods(ctx, small_set):
    """initailizeImage (MutoMLSD.cpp:226-266), through the one-shot entry
    point with no iteration: method 2 is the constant; method 1 draws
    N(2, 0.5) * 1e-7 floored at 1e-10 (the reference seeds it from
    std::random_device, here a fixed seed, so runs repeat); method 0 with no
    image given falls back to method 1; method 0 with an image starts from
    it."""
    g, rays, _ = small_set
    rays = rays[:2000]
    out = {}
    for m in (0, 1, 2):
        cfg = _em_cfg(g, num_iteration=0, initialize_method=m, initialize_value=3e-7)
        img, cost, c = ctx.mlsd_reconstruct(cfg, rays)
        assert c["iterations_done"] == 0 and len(cost) == 0
        out[m] = img
    assert np.all(out[2] == 3e-7)
    r = out[1]
    assert r.min() >= 1e-10 and abs(r.mean() / 2e-7 - 1) < 0.01 and abs(r.std() / 0.5e-7 - 1) < 0.02
    assert np.array_equal(out[0], out[1])
    img2, _, _ = ctx.mlsd_reconstruct(_em_cfg(g, num_iteration=0, initialize_method=1), rays)
    assert np.array_equal(img2, out[1])
    given = np.linspace(1e-8, 1e-6, g.n_vox)
    img3, _, _ = ctx.mlsd_reconstruct(_em_cfg(g, num_iteration=0, initialize_method=0), rays, init=given)
    assert np.array_equal(img3, given)


def _lam_field(n, seed=7):
    rng = np.random.default_rng(seed)
    return 1e-7 * (0.5 + rng.random(n))


@pytest.mark.parametrize("nb,obj,p", [(0, 0, 1.0), (1, 1, 2.0)])
def test_map_objective_fp64_matches_oracle(ctx, small_set, nb, obj, p):
    """The MAP objective operator (muto_map_forward/backward), FP64, from
    the stored lists and with the paths walked again: Phi and grad Phi equal
    the oracle's (whose gradient is pinned by central differences and by
    tangency with the reference's MAP surrogate) to 1e-10 relative (FP64
    atomics and the recomputed tails change the order of sums)."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=p, delta=0.1,
                  all_neighbours=nb, add_object_prior=obj)
    em = oracle.Em(cfg, rays)
    lam = _lam_field(g.n_vox)
    phi_ref = em.phi(lam, cfg)
    g_ref = em.grad(lam, cfg)
    plan = ctx.em_plan(cfg, rays=rays)
    for recompute in (False, True):
        sv = plan.map_forward(lam, M.MUTO_FP64, recompute=recompute)
        assert abs(sv.phi - phi_ref) <= 1e-10 * abs(phi_ref), (recompute, sv.phi, phi_ref)
        gr = sv.backward(1.0)
        np.testing.assert_allclose(gr, g_ref, rtol=1e-9, atol=1e-10 * np.abs(g_ref).max())
        gr2 = sv.backward(-2.5)  # the saved context is reusable and linear in grad_phi
        np.testing.assert_allclose(gr2, -2.5 * gr, rtol=1e-11, atol=1e-12 * np.abs(gr).max())
        sv.close()
    plan.close()


@pytest.mark.parametrize("nb,p", [(0, 2.0), (1, 1.2)])
def test_map_objective_fp64_stream_matches_oracle(ctx, small_set, nb, p):
    """The MAP objective operator on the FP64 brick stream (MUTO_FP64_STREAM,
    or MUTO_FP64 with MUTO_OPT_FP64_KERNEL 2): the lists' exact FP64 W_ij,
    sums per segment and per brick, so Phi within 1e-10 of the oracle and
    grad Phi within 1e-9; both selections give the same Phi bit for bit and
    the same gradient up to the order of the shared bricks' additions; the
    stream with the paths walked again is refused."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=p, delta=0.1,
                  all_neighbours=nb, add_object_prior=1)
    em = oracle.Em(cfg, rays)
    lam = _lam_field(g.n_vox)
    phi_ref = em.phi(lam, cfg)
    g_ref = em.grad(lam, cfg)
    plan = ctx.em_plan(cfg, rays=rays)
    sv = plan.map_forward(lam, M.MUTO_FP64_STREAM)
    assert abs(sv.phi - phi_ref) <= 1e-10 * abs(phi_ref), (sv.phi, phi_ref)
    gr = sv.backward(1.0)
    scale = np.abs(g_ref).max()
    print(f"fp64 stream map: phi rel {abs(sv.phi - phi_ref) / abs(phi_ref):.3e} "
          f"grad {np.abs(gr - g_ref).max() / scale:.3e}", flush=True)
    np.testing.assert_allclose(gr, g_ref, rtol=1e-9, atol=1e-10 * scale)
    np.testing.assert_allclose(sv.backward(-2.5), -2.5 * gr, rtol=1e-11, atol=1e-12 * scale)
    plan.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
    s2 = plan.map_forward(lam, M.MUTO_FP64)
    assert s2.phi == sv.phi
    np.testing.assert_allclose(s2.backward(1.0), gr, rtol=0, atol=1e-13 * scale)
    with pytest.raises(M.MutoError):
        plan.map_forward(lam, M.MUTO_FP64_STREAM, recompute=True)
    s2.close()
    sv.close()
    plan.close()


@pytest.fixture(scope="module")
def fd32_set():
    return W.muons("fd32")


@pytest.mark.parametrize("recompute", [False, True])
def test_map_gradient_central_differences_gpu(ctx, fd32_set, recompute):
    """grad Phi from the backward call against FP64 central differences of
    the forward call, h = 1e-6 lambda_j, on the 32^3 x 5 mm grid with 1e4
    muons, on voxels with and without entries (Phi's sums are block-ordered,
    so the forward value is the same in every run)."""
    g, rays, _ = fd32_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=0.5, p=1.5, delta=0.2,
                  all_neighbours=1, add_object_prior=1)
    plan = ctx.em_plan(cfg, rays=rays)
    lam = _lam_field(g.n_vox, 11)
    sv = plan.map_forward(lam, M.MUTO_FP64, recompute=recompute)
    gr = sv.backward(1.0)
    again = plan.map_forward(lam, M.MUTO_FP64, recompute=recompute)
    assert again.phi == sv.phi
    sv.close()
    again.close()
    mj = plan.voxel_counts()
    rng = np.random.default_rng(3)
    hit = np.nonzero(mj > 0)[0]
    miss = np.nonzero(mj == 0)[0]
    picks = list(rng.choice(hit, 8, replace=False)) + list(rng.choice(miss, 2, replace=False))
    for j in picks:
        h = 1e-6 * lam[j]
        lp, lm = lam.copy(), lam.copy()
        lp[j] += h
        lm[j] -= h
        sp = plan.map_forward(lp, M.MUTO_FP64, recompute=recompute)
        sm = plan.map_forward(lm, M.MUTO_FP64, recompute=recompute)
        fd = (sp.phi - sm.phi) / (2 * h)
        sp.close()
        sm.close()
        assert abs(fd - gr[j]) <= 1e-5 * abs(gr[j]) + 1e-7 * np.abs(gr).max(), (j, mj[j], fd, gr[j])
    plan.close()


def test_map_objective_fp32c_close_to_fp64(ctx, small_set):
    """FP32C objective operator against FP64: Phi within 1e-6 relative, the
    gradient within 1e-4 of its largest component."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=1.2, delta=0.1)
    plan = ctx.em_plan(cfg, rays=rays)
    lam = _lam_field(g.n_vox, 5)
    s64 = plan.map_forward(lam, M.MUTO_FP64)
    s32 = plan.map_forward(lam, M.MUTO_FP32C)
    g64, g32 = s64.backward(1.0), s32.backward(1.0)
    print(f"fp32c map: phi rel {abs(s32.phi - s64.phi) / abs(s64.phi):.3e} "
          f"grad {np.abs(g32 - g64).max() / np.abs(g64).max():.3e}", flush=True)
    assert abs(s32.phi - s64.phi) <= 1e-6 * abs(s64.phi)
    assert np.abs(g32 - g64).max() <= 1e-4 * np.abs(g64).max()
    # FP32C with the paths walked again (FP32 entry terms in the backward)
    r32 = plan.map_forward(lam, M.MUTO_FP32C, recompute=True)
    gr32 = r32.backward(1.0)
    print(f"fp32c recompute: grad {np.abs(gr32 - g64).max() / np.abs(g64).max():.3e}", flush=True)
    assert r32.phi == plan.map_forward(lam, M.MUTO_FP64, recompute=True).phi
    assert np.abs(gr32 - g64).max() <= 1e-4 * np.abs(g64).max()
    r32.close()
    s64.close()
    s32.close()
    plan.close()


@pytest.mark.parametrize("prec,recompute", [(M.MUTO_FP64, False), (M.MUTO_FP64, True),
                                            (M.MUTO_FP32C, False)])
def test_device_buffers_match_host_buffers(ctx, small_set, prec, recompute):
    """Device-pointer mode of the image hand-over (muto_em_plan_set_image_ex
    / get_image_ex, MUTO_MEM_DEVICE) and of the MAP operator
    (MUTO_MAP_DEVICE_BUFFERS): the same image bits as the host calls, Phi
    bit-identical, and the gradient equal up to the order of its FP64
    atomic additions (1e-13 of its largest component)."""
    from muontomography_amd.devmem import DeviceArray
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=1.2, delta=0.1,
                  all_neighbours=1)
    plan = ctx.em_plan(cfg, rays=rays)
    lam = _lam_field(g.n_vox, 9)
    tl = DeviceArray.from_numpy(lam)
    plan.set_image(tl)
    out = DeviceArray(g.n_vox)
    out.fill_bytes(0xFF)
    plan.image(out=out)
    host = plan.image()
    assert np.array_equal(out.numpy().view(np.uint64), lam.view(np.uint64))
    assert np.array_equal(host.view(np.uint64), lam.view(np.uint64))
    if not recompute:
        plan.run(2, prec)
        plan.image(out=out)
        assert np.array_equal(out.numpy().view(np.uint64), plan.image().view(np.uint64))
    sh = plan.map_forward(lam, prec, recompute=recompute)
    sd = plan.map_forward(tl, prec, recompute=recompute)
    assert sd.phi == sh.phi, (sd.phi, sh.phi)
    gh = sh.backward(1.0)
    gd = sd.backward(1.0)
    assert isinstance(gd, DeviceArray)
    gd = gd.numpy()
    scale = np.abs(gh).max()
    assert scale > 0
    np.testing.assert_allclose(gd, gh, rtol=0, atol=1e-13 * scale)
    # the caller's buffer is overwritten (not accumulated into) on a second call
    o2 = DeviceArray.from_numpy(np.full(g.n_vox, 7.0))
    sd.backward(-2.0, out=o2)
    np.testing.assert_allclose(o2.numpy(), -2.0 * gh, rtol=0, atol=1e-12 * scale)
    sh.close()
    sd.close()
    plan.close()


def test_plan_without_lists_recomputes(ctx, small_set):
    """A plan over its list budget keeps no path lists: the same per-voxel
    counts from a walk, EM runs refused, and the MAP operator (paths walked
    again, FP64 and FP32C) equal to the list plan's recomputed one."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=2.0, delta=0.1)
    full = ctx.em_plan(cfg, rays=rays)
    lean = ctx.em_plan(cfg, rays=rays, list_budget=0)
    assert lean.counters()["nnz"] == full.counters()["nnz"]
    np.testing.assert_array_equal(lean.voxel_counts(), full.voxel_counts())
    lean.set_image(np.full(g.n_vox, 1e-7))
    with pytest.raises(M.MutoError):
        lean.run(1, M.MUTO_FP64)
    lam = _lam_field(g.n_vox, 9)
    a = full.map_forward(lam, M.MUTO_FP64, recompute=True)
    b = lean.map_forward(lam, M.MUTO_FP64)
    assert abs(a.phi - b.phi) <= 1e-12 * abs(a.phi)
    np.testing.assert_allclose(b.backward(1.0), a.backward(1.0), rtol=1e-11,
                               atol=1e-12 * np.abs(a.backward(1.0)).max())
    # FP32C on a plan without lists walks the paths again too (FP32 terms
    # in its backward)
    c = lean.map_forward(lam, M.MUTO_FP32C)
    assert c.phi == b.phi
    gb = b.backward(1.0)
    assert np.abs(c.backward(1.0) - gb).max() <= 1e-4 * np.abs(gb).max()
    c.close()
    a.close()
    b.close()
    full.close()
    lean.close()


@pytest.mark.parametrize("straight,shape", [(0, None), (1, None), (0, (7, 33, 18))])
def test_map_recompute_tiles_match_per_entry_atomics(ctx, small_set, straight, shape):
    """The MAP operator with the paths walked again per 16^3 tile (the tile
    records fit the list budget of a plan without lists: moments per run in
    the forward pass, sums in LDS in the backward) against one lane per
    track and one atomic per entry (a budget of 0 keeps no records): the
    same entry terms from walk_path's voxels, lengths and tails, so Phi
    agrees to 1e-12 relative and the FP64 gradients up to the order of the
    sums (1e-12 of the largest component), the FP32C ones within 1e-5 (FP32
    against FP64 sums of the same FP32 terms); kinked and straight paths on
    a 40^3 grid (partial tiles at the far sides) and on a 7x33x18 grid over
    the same cube (one axis shorter than a tile, unequal voxel sides), and
    the FP32C gradient within 1e-4 of the FP64 one."""
    g, rays, _ = small_set
    if shape is not None:
        ext = g.dx * g.nx
        g = make_grid(shape, tuple(ext / k for k in shape), g.x_min)
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, beta=1.0, p=2.0, delta=0.1,
                  straight_line=straight)
    probe = ctx.em_plan(cfg, rays=rays)
    nnz = probe.counters()["nnz"]
    probe.close()
    per_entry = ctx.em_plan(cfg, rays=rays, list_budget=0)
    tiled = ctx.em_plan(cfg, rays=rays, list_budget=20 * nnz - 1)
    lam = _lam_field(g.n_vox, 4)
    grads = {}
    for prec, tol in ((M.MUTO_FP64, 1e-12), (M.MUTO_FP32C, 1e-5)):
        a = per_entry.map_forward(lam, prec)
        b = tiled.map_forward(lam, prec)
        assert abs(a.phi - b.phi) <= 1e-12 * abs(a.phi), (a.phi, b.phi)
        c = tiled.map_forward(lam, prec)
        assert c.phi == b.phi  # the same sums in the same order
        c.close()
        ga, gb = a.backward(1.0), b.backward(1.0)
        scale = np.abs(ga).max()
        assert scale > 0
        err = np.abs(gb - ga).max() / scale
        print(f"straight {straight} precision {prec}: tiled against per-entry {err:.3e}", flush=True)
        assert err <= tol, err
        # a second call on the same records gives the same sums up to their order
        np.testing.assert_allclose(b.backward(1.0), gb, rtol=0, atol=1e-13 * scale)
        grads[prec] = gb
        a.close()
        b.close()
    assert per_entry.counters()["n_map_tile_records"] == 0
    n_rec = tiled.counters()["n_map_tile_records"]
    assert 0 < n_rec < nnz, (n_rec, nnz)
    g64 = grads[M.MUTO_FP64]
    assert np.abs(grads[M.MUTO_FP32C] - g64).max() <= 1e-4 * np.abs(g64).max()
    per_entry.close()
    tiled.close()


def _records_from_rays(rays, zl):
    """Float32 tracker records (e_in, (x, y) per layer, e_out) of the rays'
    in and out lines evaluated at the layer heights (3 above, 3 below)."""
    n = rays.shape[0]
    rec = np.empty((n, 2 * len(zl) + 2), np.float32)
    rec[:, 0] = rays["ein"]
    rec[:, -1] = rays["eout"]
    for i, z in enumerate(zl):
        p, d = (rays["pin"], rays["din"]) if i < 3 else (rays["pout"], rays["dout"])
        t = (z - p[:, 2]) / d[:, 2]
        rec[:, 1 + 2 * i] = p[:, 0] + t * d[:, 0]
        rec[:, 2 + 2 * i] = p[:, 1] + t * d[:, 1]
    return rec


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_rays_from_records_bitexact(ctx, small_set, dt):
    """Records to rays on the GPU (NaN filter, per-side principal-axis fits)
    equal the oracle's bit for bit, NaN records dropped in order."""
    g, rays, _ = small_set
    zi, zo = float(rays["pin"][0, 2]), float(rays["pout"][0, 2])
    zl = [zi + 1000.0, zi + 500.0, zi, zo, zo - 500.0, zo - 1000.0]
    rec = W.records_from_rays(rays, zl, sigma=0.5, seed=2, dtype=dt)
    rng = np.random.default_rng(8)
    bad = rng.choice(rec.shape[0], 300, replace=False)
    rec[bad, rng.integers(0, rec.shape[1], 300)] = np.nan
    rg = ctx.rays_from_records(rec, zl, 3)
    ro = oracle.rays_from_records(rec, zl, 3)
    assert rg.shape[0] == ro.shape[0] == rec.shape[0] - 300
    assert np.array_equal(rg.view(np.uint8), ro.view(np.uint8))
    assert ctx.kernel_ms() > 0.0


def test_cli_poca_then_mlsd_matches_library(ctx, small_set, tmp_path):
    """The reference-style driver (muto_cli: JSON config, record files,
    image files) end to end: PoCA, then MLSD started from the PoCA image
    file, equal to the library called on the rays the C++ reader parsed
    (images compared as the float32 values the files hold)."""
    import json
    import subprocess
    from muontomography_amd.imagefile import read_image
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    g, rays, _ = small_set
    rays = rays[:20000]
    zi, zo = float(rays["pin"][0, 2]), float(rays["pout"][0, 2])
    zl = [zi + 1000.0, zi + 500.0, zi, zo, zo - 500.0, zo - 1000.0]
    recf = tmp_path / "muons.bin"
    recf.write_bytes(_records_from_rays(rays, zl).tobytes())
    rf = tmp_path / "rays.bin"
    r = subprocess.run([os.path.join(root, "tests", "cxx", "dropin_check"), "rays", str(rf), "4", "3"]
                       + [repr(z) for z in zl] + ["--", str(recf)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rc = np.frombuffer(rf.read_bytes(), M.RAY_DTYPE)
    assert rc.shape[0] == rays.shape[0]
    grid = dict(nx=g.nx, ny=g.ny, nz=g.nz, x_min=g.x_min, y_min=g.y_min, z_min=g.z_min,
                dx=g.dx, dy=g.dy, dz=g.dz)
    common = dict(data_files=[str(recf)], detector_height=zl, detector_middle=3)
    poca_out = tmp_path / "poca.bin"
    cfgp = dict(common, output_file=str(poca_out),
                poca=dict(grid=grid, length_weighted=True, straight_line=False, eps=1e-8))
    (tmp_path / "poca.json").write_text(json.dumps(cfgp))
    cli = os.path.join(root, "muontomography_amd", "muto_cli")
    r = subprocess.run([cli, str(tmp_path / "poca.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    pimg, ph = read_image(poca_out)
    assert ph["method"] == "poca" and ph["n_ray"] == rays.shape[0] and ph["nx"] == g.nx
    ref, _ = ctx.poca_reconstruct(poca_cfg(g, length_weighted=1), rc)
    np.testing.assert_allclose(pimg, ref.astype(np.float32), rtol=1e-6, atol=0)
    mlsd_out = tmp_path / "mlsd.bin"
    mk = dict(grid=grid, update_with_mean=True, mean_percentile=100, num_iteration=5,
              initialize_method=0, initial_image_file=str(poca_out), error_xy=0.5,
              error_dz_outer=500.0, error_dz_inner=500.0)
    cfgm = dict(common, output_file=str(mlsd_out), mlsd=mk)
    (tmp_path / "mlsd.json").write_text(json.dumps(cfgm))
    r = subprocess.run([cli, str(tmp_path / "mlsd.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    mimg, mh = read_image(mlsd_out)
    assert mh["method"] == "mlsd" and mh["n_iter"] == 5
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100, num_iteration=5, initialize_method=0)
    ref, _, _ = ctx.mlsd_reconstruct(cfg, rc, init=pimg)
    np.testing.assert_allclose(mimg, ref.astype(np.float32), rtol=1e-5, atol=0)
    # the median (the reference's default rule) on the FP64 brick stream,
    # through the drop-in's "precision" key, against the reference-order run
    mk2 = dict(mk, update_with_mean=False, precision="fp64_stream")
    (tmp_path / "mlsd2.json").write_text(json.dumps(dict(common, output_file=str(mlsd_out), mlsd=mk2)))
    r = subprocess.run([cli, str(tmp_path / "mlsd2.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    mimg2, _ = read_image(mlsd_out)
    cfg2 = _em_cfg(g, update_with_mean=0, num_iteration=5, initialize_method=0)
    ref2, _, _ = ctx.mlsd_reconstruct(cfg2, rc, init=pimg)
    np.testing.assert_allclose(mimg2, ref2.astype(np.float32), rtol=1e-5, atol=0)
    # the other Siddon consumers the reference's driver dispatches
    for method, extra, fn in (("pocut", dict(cut_percentile=0.8),
                               lambda c: ctx.pocut_reconstruct(c, rc, 0.8)),
                              ("slr", {}, lambda c: ctx.slr_reconstruct(c, rc))):
        out = tmp_path / f"{method}.bin"
        cfgx = dict(common, output_file=str(out), **{method: dict(grid=grid, length_weighted=True, **extra)})
        (tmp_path / f"{method}.json").write_text(json.dumps(cfgx))
        r = subprocess.run([cli, str(tmp_path / f"{method}.json")], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        ximg, xh = read_image(out)
        assert xh["method"] == method and xh["n_iter"] == 1
        ref, _ = fn(poca_cfg(g, length_weighted=1))
        np.testing.assert_allclose(ximg, ref.astype(np.float32), rtol=1e-6, atol=0)


def test_batch_of_scans_matches_each_scan_alone(ctx):
    """Six scans of different grids, voxel sizes and track counts in one
    batch (one launch per pass, one graph per iteration): each scan's images
    and costs equal those of its plan run alone in FP32C up to the order of
    FP64 atomics and the fixed-point scale of the backprojection sums, which
    the batch sets from the largest |c_ij| of all scans: costs rtol 1e-7,
    images within 1e-5 of their largest value after 5 iterations (FP32C's
    own distance to FP64 is larger)."""
    specs = W.scan_batch(n_scans=6, seed=5, n_scale=0.05)
    plans, alone = [], []
    for g, gcfg, nt in specs:
        rays, _ = M.generate(gcfg, nt)
        cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
        a = ctx.em_plan(cfg, rays=rays)
        b = ctx.em_plan(cfg, rays=rays)
        img = np.full(g.n_vox, 1e-7)
        a.set_image(img)
        b.set_image(img)
        plans.append(a)
        alone.append(b)
    batch = ctx.batch(plans)
    cost = batch.run(5)
    ms, c = batch.stats()
    assert ms > 0 and c["nnz"] == sum(p.counters()["nnz"] for p in plans)
    for s, (a, b) in enumerate(zip(plans, alone)):
        cb = b.run(5, M.MUTO_FP32C)
        np.testing.assert_allclose(cost[:, s], cb, rtol=1e-7)
        ia, ib = a.image(), b.image()
        m = ib > 0
        assert np.abs(ia - ib).max() <= 1e-5 * np.abs(ib).max(), (s, np.abs(ia - ib).max())
        assert np.all(ia[~m] == 0)
    # a second run continues from the images written back
    cost2 = batch.run(2)
    for s, b in enumerate(alone):
        np.testing.assert_allclose(cost2[:, s], b.run(2, M.MUTO_FP32C), rtol=1e-7)
    batch.close()
    for p in plans + alone:
        p.close()


@pytest.mark.parametrize("shape", [(16, 16, 16, 256), (16, 8, 32, 256), (16, 16, 32, 512),
                                   (16, 16, 64, 1024), (32, 16, 32, 1024)])
def test_fp32c_block_sizes(ctx, small_set, shape):
    """The FP32C passes with 256-, 512- and 1024-thread blocks (chunks of
    4096, 8192 and 16384 entries) and their brick shapes: 10 iterations
    within 1e-4 of FP64, costs within 1e-6."""
    g, rays, _ = small_set
    cfg = _em_cfg(g, update_with_mean=1, mean_percentile=100)
    plan = ctx.em_plan(cfg, rays=rays)
    plan.set_option(M.MUTO_OPT_BRICK, M.brick_code(shape))
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    c32 = plan.run(10, M.MUTO_FP32C)
    a32 = plan.image()
    assert plan.counters()["n_segments"] > 0
    plan.set_image(img)
    c64 = plan.run(10, M.MUTO_FP64)
    mx, rms, _ = _rel_stats(a32, plan.image(), plan.voxel_counts())
    assert mx <= 1e-4, (shape, mx, rms)
    np.testing.assert_allclose(c32, c64, rtol=1e-6)
    plan.close()


@pytest.mark.parametrize("rule", ["mean", "median"])
def test_autotune_keeps_image_and_picks_a_candidate(ctx, small_set, rule):
    """The brick-shape autotuner (timing the plan's own update rule: the mean,
    or the median by brackets) leaves the image as it was, picks one of the
    candidates it timed, and the FP32C run with the chosen shape matches the
    same shape set by the option (and stays within 1e-4 of FP64)."""
    g, rays, _ = small_set
    kw = dict(update_with_mean=1, mean_percentile=100) if rule == "mean" else dict(update_with_mean=0)
    cfg = _em_cfg(g, **kw)
    plan = ctx.em_plan(cfg, rays=rays)
    img = np.full(g.n_vox, 1e-7)
    plan.set_image(img)
    shapes = [(16, 16, 16, 256), (8, 8, 32, 256), (16, 16, 32, 512), (16, 16, 64, 1024)]
    best, ms = plan.autotune(iters=2, shapes=shapes)
    assert best in shapes and all(np.isfinite(v) and v > 0 for v in ms.values()), (best, ms)
    np.testing.assert_array_equal(plan.image(), img)
    c_t = plan.run(10, M.MUTO_FP32C)
    a_t = plan.image()
    other = ctx.em_plan(cfg, rays=rays)
    other.set_option(M.MUTO_OPT_BRICK, M.brick_code(best))
    other.set_image(img)
    c_o = other.run(10, M.MUTO_FP32C)
    np.testing.assert_allclose(c_t, c_o, rtol=1e-9)
    np.testing.assert_allclose(a_t, other.image(), rtol=1e-6, atol=1e-6 * np.abs(a_t).max())
    other.set_image(img)
    other.run(10, M.MUTO_FP64)
    mj = other.voxel_counts()
    mx, _, _ = _rel_stats(a_t, other.image(), mj)
    if rule == "mean":
        assert mx <= 1e-4, mx
    else:  # as test_fp32c_order_statistics_match_fp64: few-track voxels within 5e-4
        mx8, _, _ = _rel_stats(a_t, other.image(), np.where(mj >= 8, mj, 0))
        assert mx <= 5e-4 and mx8 <= 1e-4, (mx, mx8)
    plan.close()
    other.close()
