"""Synthetic workloads: grids, block layouts and muon sets of the benchmark
configurations (sizes in mm, momenta in GeV/c).

  small     40^3 x 25 mm, 1e5 muons, one iron and one lead 100 mm cube
  siddon200 200^3 x 10 mm, 1e7 muons, iron/lead/water blocks 50-300 mm
  em256     256^3 x 10 mm, 1e7 muons whose chords cross 80-600 voxels
  fd32      32^3 x 5 mm, 1e4 muons (MAP gradient finite differences)
  map128    128^3 x 15 mm, 1e6 muons (MAP gradient run)
"""
import numpy as np

from . import gen_config, gen_layout, generate, gen_truth, make_grid

X0_AIR = 303900.0
X0_IRON = 17.6
X0_LEAD = 5.6
X0_WATER = 360.0


def cube_grid(n, d):
    """n^3 voxels of d mm centred on the origin."""
    return make_grid(n, float(d), -0.5 * n * d)


def workload(name, n_muons=None, seed=None):
    """Returns (grid, gen_cfg, default n_muons)."""
    if name == "small":
        g = cube_grid(40, 25.0)
        s = 1 if seed is None else seed
        blocks = gen_layout(g, s, 2, 100.0, 100.0, [X0_IRON, X0_LEAD])
        cfg = gen_config(g, s, blocks)
        n = 100_000
    elif name == "siddon200":
        g = cube_grid(200, 10.0)
        s = 2 if seed is None else seed
        blocks = gen_layout(g, s, 8, 50.0, 300.0, [X0_IRON, X0_LEAD, X0_WATER])
        cfg = gen_config(g, s, blocks)
        n = 10_000_000
    elif name == "em256":
        g = cube_grid(256, 10.0)
        s = 3 if seed is None else seed
        blocks = gen_layout(g, s, 12, 50.0, 300.0, [X0_IRON, X0_LEAD, X0_WATER])
        cfg = gen_config(g, s, blocks, min_voxels=80, max_voxels=600)
        n = 10_000_000
    elif name == "fd32":
        g = cube_grid(32, 5.0)
        s = 4 if seed is None else seed
        blocks = gen_layout(g, s, 2, 20.0, 40.0, [X0_IRON, X0_LEAD])
        cfg = gen_config(g, s, blocks)
        n = 10_000
    elif name == "map128":
        g = cube_grid(128, 15.0)
        s = 5 if seed is None else seed
        blocks = gen_layout(g, s, 8, 50.0, 300.0, [X0_IRON, X0_LEAD, X0_WATER])
        cfg = gen_config(g, s, blocks)
        n = 1_000_000
    else:
        raise KeyError(name)
    return g, cfg, (n if n_muons is None else n_muons)


def muons(name, n_muons=None, seed=None):
    g, cfg, n = workload(name, n_muons, seed)
    rays, tried = generate(cfg, n)
    return g, rays, {"tried": tried, "acceptance": n / max(tried, 1), "truth": gen_truth(cfg)}


def scan_batch(n_scans=64, seed=77, n_scale=1.0):
    """Many independent scans: cubic grids 32..128 voxels, 10-30 mm voxels,
    track counts log-uniform in [1e4, 1e6] (times n_scale)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scans):
        n = int(rng.integers(32, 129))
        d = float(rng.uniform(10.0, 30.0))
        nt = int(round(10 ** rng.uniform(4, 6) * n_scale))
        g = cube_grid(n, d)
        blocks = gen_layout(g, 1000 + k, 4, 3 * d, 12 * d, [X0_IRON, X0_LEAD, X0_WATER])
        out.append((g, gen_config(g, 1000 + k, blocks), max(nt, 100)))
    return out


def records_from_rays(rays, zlayers, n_middle=3, sigma=0.0, seed=0, dtype=np.float32):
    """Tracker records (e_in, (x, y) per plane, e_out) of the rays' in and out
    lines evaluated at the plane heights, hits smeared by sigma mm: the input
    format of MutoFile::getMuonData (src/MutoFile.cpp:19-82)."""
    n = rays.shape[0]
    zl = np.asarray(zlayers, np.float64)
    rec = np.empty((n, 2 * zl.shape[0] + 2), dtype)
    rec[:, 0] = rays["ein"]
    rec[:, -1] = rays["eout"]
    rng = np.random.default_rng(seed)
    for i, z in enumerate(zl):
        p, d = (rays["pin"], rays["din"]) if i < n_middle else (rays["pout"], rays["dout"])
        t = (z - p[:, 2]) / d[:, 2]
        x = p[:, 0] + t * d[:, 0]
        y = p[:, 1] + t * d[:, 1]
        if sigma > 0:
            x = x + rng.normal(0.0, sigma, n)
            y = y + rng.normal(0.0, sigma, n)
        rec[:, 1 + 2 * i] = x
        rec[:, 2 + 2 * i] = y
    return rec
