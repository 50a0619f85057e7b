"""Constructed MLSD inputs whose per-track data follow in closed form from
the geometry: kinked tracks (an incoming and an outgoing line through a known
kink point K), degenerate rays whose PoCA point is exactly a chosen point,
and the expected selection, angles, displacements and merged two-leg paths
of src/MutoMLSD.cpp:268-415.  Shared by the oracle tests (CPU) and the GPU
parity tests, so that both are held to the same closed forms."""
import numpy as np

import muontomography_amd as M
from muontomography_amd._abi import make_grid


def ray(pin, din, pout, dout, e=3.0):
    r = np.zeros(1, M.RAY_DTYPE)
    r["pin"], r["din"], r["pout"], r["dout"] = pin, din, pout, dout
    r["ein"] = r["eout"] = e
    return r


def kinked(K, din, dout, z_top=1000.0, z_bot=-1000.0):
    """Track through the kink point K: pin on the incoming line at z_top,
    pout on the outgoing line at z_bot (directions point up, d.z > 0, as the
    reference's record reader leaves them)."""
    K = np.asarray(K, float)
    din = np.asarray(din, float)
    dout = np.asarray(dout, float)
    pin = K + (z_top - K[2]) / din[2] * din
    pout = K + (z_bot - K[2]) / dout[2] * dout
    return ray(pin, din, pout, dout)


def kink_closed_form(K, din, dout, z_min):
    """(Sx, Sy, Dx, Dy) of a track kinked at K.  Both lines pass through K,
    so where they cross the plane z = z_min they are (z_min - K.z) tan a
    apart per plane, and the reference's projection factors
    cos(a_in) Lxy cos(a_out) / cos(a_out - a_in) (src/MutoMLSD.cpp:356-376)
    reduce, by tan a - tan b = sin(a - b) / (cos a cos b), to
        D = (z_min - K.z) tan(a_out - a_in) |din| / din.z
    in each plane, with Lxy = |din| / din.z.  S = a_in - a_out."""
    din = np.asarray(din, float)
    dout = np.asarray(dout, float)
    axi, axo = np.arctan(din[0] / din[2]), np.arctan(dout[0] / dout[2])
    ayi, ayo = np.arctan(din[1] / din[2]), np.arctan(dout[1] / dout[2])
    lxy = np.linalg.norm(din) / din[2]
    h = z_min - K[2]
    return np.array([axi - axo, ayi - ayo, h * np.tan(axo - axi) * lxy, h * np.tan(ayo - ayi) * lxy])


def random_kinks(g, n, seed, in_plane=None):
    """n kinked tracks with kink points well inside grid g, incoming slopes up
    to 0.5 and kinks up to 0.05 rad.  in_plane = "xz" / "yz" keeps every
    track in one plane (y or x components zero)."""
    rng = np.random.default_rng(seed)
    lo = np.array([g.x_min, g.y_min, g.z_min])
    hi = lo + np.array([g.dx * g.nx, g.dy * g.ny, g.dz * g.nz])
    rays, want = [], []
    for _ in range(n):
        K = lo + (hi - lo) * rng.uniform(0.2, 0.8, 3)
        si = rng.uniform(-0.5, 0.5, 2)
        so = si + rng.uniform(-0.05, 0.05, 2)
        if in_plane == "xz":
            si[1] = so[1] = 0.0
        elif in_plane == "yz":
            si[0] = so[0] = 0.0
        din = np.array([si[0], si[1], 1.0])
        dout = np.array([so[0], so[1], 1.0])
        rays.append(kinked(K, din, dout))
        want.append(kink_closed_form(K, din, dout, g.z_min))
    return np.concatenate(rays), np.array(want)


def window_rays(g, z_det, eps=1e-8):
    """Degenerate rays (pin = pout = Q, din = dout): the PoCA lines are
    parallel, so the PoCA point is 0.5 (Q + Q) = Q exactly.  Ray 0 fixes the
    detector window (zDetMin = rays[0].pout.z, zDetMax = rays[0].pin.z,
    src/MutoMLSD.cpp:273-274).  Every other ray puts one coordinate of Q on a
    boundary of the selection window of src/MutoMLSD.cpp:298-301, or one ulp
    outside it:
        x_min + eps <= P.x <= x_min + dx nx - eps   (and y, z likewise)
        zDetMin + eps <= P.z < zDetMax - eps.
    Returns the rays and the accept flag each must get."""
    zb, zt = z_det
    lo = np.array([g.x_min, g.y_min, g.z_min])
    ext = np.array([g.dx * g.nx, g.dy * g.ny, g.dz * g.nz])
    c = lo + 0.5 * ext
    c[2] = 0.5 * (max(zb, lo[2]) + min(zt, lo[2] + ext[2]))
    up = np.array([0.0, 0.0, 1.0])
    rays = [ray([c[0], c[1], zt], up, [c[0], c[1], zb], up)]
    want = [True]

    def at(axis, v, ok):
        Q = c.copy()
        Q[axis] = v
        rays.append(ray(Q, up, Q, up))
        want.append(ok)

    for a in range(3):
        # the same double expressions as the reference's test
        b_lo = lo[a] + eps
        b_hi = lo[a] + [g.dx, g.dy, g.dz][a] * [g.nx, g.ny, g.nz][a] - eps
        if a == 2:
            d_lo, d_hi = zb + eps, zt - eps
            # which window binds at each end: the grid's or the detector's
            if d_lo <= b_lo:
                at(2, b_lo, True); at(2, np.nextafter(b_lo, -np.inf), False)
            else:
                at(2, d_lo, True); at(2, np.nextafter(d_lo, -np.inf), False)
            if d_hi <= b_hi:  # detector top: zDetMax - eps <= P.z rejects
                at(2, d_hi, False); at(2, np.nextafter(d_hi, -np.inf), True)
            else:
                at(2, b_hi, True); at(2, np.nextafter(b_hi, np.inf), False)
            continue
        at(a, b_lo, True); at(a, np.nextafter(b_lo, -np.inf), False)
        at(a, b_hi, True); at(a, np.nextafter(b_hi, np.inf), False)
    return np.concatenate(rays), np.array(want)


# ---- two-leg merge (src/MutoMLSD.cpp:382-415) on a 4^3 grid of 10 mm voxels
MERGE_GRID = make_grid(4, 10.0, -20.0)
MERGE_THETA = 1e-3


def _j(x, y, z):
    return x + 4 * y + 16 * z


def merge_cases():
    """Vertical incoming track at (x, y) = (4, 3) (voxel column (2, 2)),
    kinked at K by MERGE_THETA toward x, so the outgoing leg stays in the
    column (it drifts 0.023 mm over the 23 mm below K).  Layers are
    [-20,-10), [-10,0), [0,10), [10,20).

    "inside": K.z = 3, in layer 2.  Siddon gives the in-leg (layer 3: 10,
    layer 2: 20 - 10 - 3 = 7) and the out-leg (layer 2: 3/c, layer 1: 10/c,
    layer 0: 10/c), c = cos(theta).  The merge adds out[0] into in.back() and
    drops out[0]: lengths (10, 7 + 3/c, 10/c, 10/c).

    "empty_in": K.z = 20 - 1e-3, just under the top face, with pin 1e5 mm
    above.  The in-leg's alpha range inside the grid, 1e-3 / 99980, is below
    Siddon's alpha eps (1e-6), so the in-leg is empty and out[0] (the top
    layer's 1e-3/c) is dropped: lengths (10/c, 10/c, 10/c) in layers 2, 1, 0.
    The out-leg ends 10 mm under the grid, so its start is 2e-5 in alpha
    from the top face: above the eps, so Siddon's near-face start rule
    (src/MutoSiddon.cpp:85-103) does not apply.

    Returns {name: (rays, vox, len, tail, n_in, n_out)} with the tails T_k =
    sum of the lengths after k and the two legs' Siddon entry counts."""
    c = np.cos(MERGE_THETA)
    din = np.array([0.0, 0.0, 1.0])
    dout = np.array([np.sin(MERGE_THETA), 0.0, np.cos(MERGE_THETA)])
    out = {}
    for name, kz, zt, zb, vox, ln, n_io in (
            ("inside", 3.0, 1000.0, -980.0, [_j(2, 2, 3), _j(2, 2, 2), _j(2, 2, 1), _j(2, 2, 0)],
             [10.0, 7.0 + 3.0 / c, 10.0 / c, 10.0 / c], (2, 3)),
            ("empty_in", 20.0 - 1e-3, 1e5, -30.0, [_j(2, 2, 2), _j(2, 2, 1), _j(2, 2, 0)],
             [10.0 / c, 10.0 / c, 10.0 / c], (0, 4))):
        r = kinked([4.0, 3.0, kz], din, dout, z_top=zt, z_bot=zb)
        ln = np.array(ln)
        tail = np.concatenate([np.cumsum(ln[::-1])[::-1][1:], [0.0]])
        out[name] = (r, np.array(vox, np.uint32), ln, tail) + n_io
    return out
