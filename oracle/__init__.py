"""Python handle on the CPU FP64 oracle (liboracle_muto.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP kernels are compared with.
Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from muontomography_amd._abi import (RAY_DTYPE, Counters, Grid, MlsdCfg, PocaCfg, ptr)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_muto.so")

_lib = None


class OrcEm(C.Structure):
    _fields_ = [("grid", Grid), ("n_sel", C.c_size_t), ("sel", C.c_void_p), ("trk4", C.c_void_p),
                ("p", C.c_void_p), ("pscat", C.c_void_p), ("row_ptr", C.c_void_p),
                ("vox", C.c_void_p), ("len", C.c_void_p), ("tail", C.c_void_p),
                ("col_ptr", C.c_void_p), ("slot", C.c_void_p), ("S", C.c_void_p),
                ("n_empty", C.c_int64), ("n_out_of_grid", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_x86_u32.restype = C.c_uint32
        L.orc_x86_u32.argtypes = [C.c_double]
        L.orc_siddon_csr.argtypes = [C.POINTER(Grid), C.c_size_t, P, P, P, P, P, P, P]
        L.orc_siddon_digest.argtypes = [C.POINTER(Grid), C.c_size_t, P, P, P, P]
        L.orc_poca_points.argtypes = [C.POINTER(Grid), C.c_size_t, P, C.c_double, C.c_int, P, P, P,
                                      C.POINTER(Counters)]
        L.orc_poca_reconstruct.argtypes = [C.POINTER(PocaCfg), C.c_size_t, P, P, C.POINTER(Counters)]
        L.orc_pocut_reconstruct.argtypes = [C.POINTER(PocaCfg), C.c_double, C.c_size_t, P, P,
                                            C.POINTER(Counters)]
        L.orc_rays_from_records.argtypes = [C.c_size_t, C.c_int, C.c_int, P, P, C.c_int, P,
                                            C.POINTER(C.c_size_t)]
        L.orc_fit_line.argtypes = [P, P, P, C.c_int, C.c_double, P, P]
        L.orc_slr_reconstruct.argtypes = [C.POINTER(PocaCfg), C.c_size_t, P, P, C.POINTER(Counters)]
        L.orc_em_prepare.argtypes = [C.POINTER(MlsdCfg), C.c_size_t, P, C.POINTER(OrcEm),
                                     C.POINTER(Counters)]
        L.orc_em_from_prepared.argtypes = [C.POINTER(Grid), C.c_size_t, P, P, P, P, P,
                                           C.POINTER(OrcEm)]
        L.orc_em_free.argtypes = [C.POINTER(OrcEm)]
        L.orc_em_iterate.restype = C.c_double
        L.orc_em_iterate.argtypes = [C.POINTER(MlsdCfg), C.POINTER(OrcEm), P, P, P]
        L.orc_em_cij.argtypes = [C.POINTER(MlsdCfg), C.POINTER(OrcEm), P, P, P]
        L.orc_em_mj.argtypes = [C.POINTER(OrcEm), P]
        L.orc_mlsd_reconstruct.argtypes = [C.POINTER(MlsdCfg), C.c_size_t, P, P, P, P,
                                           C.POINTER(Counters)]
        L.orc_map_reconstruct.argtypes = L.orc_mlsd_reconstruct.argtypes
        L.orc_cubic_solve.restype = C.c_uint32
        L.orc_cubic_solve.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_int)]
        L.orc_map_voxel_step.argtypes = [C.POINTER(MlsdCfg), P, P, P, P, P]
        L.orc_map_surrogate.restype = C.c_double
        L.orc_map_surrogate.argtypes = [C.POINTER(MlsdCfg), C.c_uint32, C.c_double, C.c_double,
                                        C.c_double, P, P, C.c_int]
        L.orc_map_neighbours.restype = C.c_int
        L.orc_map_neighbours.argtypes = [C.POINTER(MlsdCfg), P, C.c_uint32, P, P]
        L.orc_map_phi.restype = C.c_double
        L.orc_map_phi.argtypes = [C.POINTER(MlsdCfg), C.POINTER(OrcEm), P]
        L.orc_map_grad.argtypes = [C.POINTER(MlsdCfg), C.POINTER(OrcEm), P, P]
        L.orc_map_surrogate_grad.argtypes = [C.POINTER(MlsdCfg), C.POINTER(OrcEm), P, P]
        _lib = L
    return _lib


def x86_u32(v):
    return int(lib().orc_x86_u32(float(v)))


def siddon_csr(grid, p1, p2):
    p1 = np.ascontiguousarray(p1, np.float64).reshape(-1, 3)
    p2 = np.ascontiguousarray(p2, np.float64).reshape(-1, 3)
    n = p1.shape[0]
    row = np.zeros(n + 1, np.uint64)
    assert lib().orc_siddon_csr(C.byref(grid), n, ptr(p1), ptr(p2), ptr(row), None, None, None,
                                None) == 0
    nnz = int(row[-1])
    vx, vy, vz = (np.empty(max(nnz, 1), np.uint32) for _ in range(3))
    ln = np.empty(max(nnz, 1), np.float64)
    assert lib().orc_siddon_csr(C.byref(grid), n, ptr(p1), ptr(p2), ptr(row), ptr(vx), ptr(vy),
                                ptr(vz), ptr(ln)) == 0
    return row, vx[:nnz], vy[:nnz], vz[:nnz], ln[:nnz]


def siddon_digest(grid, p1, p2):
    """Per-segment counts and list digests (OpenMP over segments)."""
    p1 = np.ascontiguousarray(p1, np.float64).reshape(-1, 3)
    p2 = np.ascontiguousarray(p2, np.float64).reshape(-1, 3)
    n = p1.shape[0]
    cnt = np.empty(max(n, 1), np.uint32)
    dig = np.empty(max(n, 1), np.uint64)
    assert lib().orc_siddon_digest(C.byref(grid), n, ptr(p1), ptr(p2), ptr(cnt), ptr(dig)) == 0
    return cnt[:n], dig[:n]


def poca_points(grid, rays, eps=1e-8, mode=0):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    n = rays.shape[0]
    P = np.empty((max(n, 1), 3))
    v = np.empty((max(n, 1), 3), np.uint32)
    a = np.empty(max(n, 1), np.uint8)
    c = Counters()
    assert lib().orc_poca_points(C.byref(grid), n, ptr(rays), eps, mode, ptr(P), ptr(v), ptr(a),
                                 C.byref(c)) == 0
    return P[:n], v[:n], a[:n].astype(bool), c.as_dict()


def poca_reconstruct(cfg, rays):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    img = np.empty(cfg.grid.n_vox)
    c = Counters()
    assert lib().orc_poca_reconstruct(C.byref(cfg), rays.shape[0], ptr(rays), ptr(img),
                                      C.byref(c)) == 0
    return img, c.as_dict()


def pocut_reconstruct(cfg, rays, cut_percentile=1.0):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    img = np.empty(cfg.grid.n_vox)
    c = Counters()
    assert lib().orc_pocut_reconstruct(C.byref(cfg), float(cut_percentile), rays.shape[0], ptr(rays),
                                       ptr(img), C.byref(c)) == 0
    return img, c.as_dict()


def slr_reconstruct(cfg, rays):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    img = np.empty(cfg.grid.n_vox)
    c = Counters()
    assert lib().orc_slr_reconstruct(C.byref(cfg), rays.shape[0], ptr(rays), ptr(img),
                                     C.byref(c)) == 0
    return img, c.as_dict()


def rays_from_records(records, zlayers, n_middle):
    rec = np.ascontiguousarray(records)
    z = np.ascontiguousarray(zlayers, np.float64)
    n = rec.shape[0]
    out = np.zeros(max(n, 1), RAY_DTYPE)
    k = C.c_size_t()
    assert lib().orc_rays_from_records(n, z.shape[0], int(n_middle), ptr(z), ptr(rec), rec.dtype.itemsize,
                                       ptr(out), C.byref(k)) == 0
    return out[:k.value]


def fit_line(x, y, z, zref):
    x, y, z = (np.ascontiguousarray(v, np.float64) for v in (x, y, z))
    pt = np.empty(3)
    d = np.empty(3)
    lib().orc_fit_line(ptr(x), ptr(y), ptr(z), x.shape[0], float(zref), ptr(pt), ptr(d))
    return pt, d


class Em:
    """Prepared MLSD problem (orc_em) with numpy views of its arrays."""

    def __init__(self, cfg, rays=None, prepared=None):
        self.cfg = cfg
        self.s = OrcEm()
        if prepared is not None:
            trk4, p, row, vox, ln = prepared
            self._k = [np.ascontiguousarray(trk4, np.float64),
                       None if p is None else np.ascontiguousarray(p, np.float64),
                       np.ascontiguousarray(row, np.uint64), np.ascontiguousarray(vox, np.uint32),
                       np.ascontiguousarray(ln, np.float64)]
            t, pp, r, v, l_ = self._k
            assert lib().orc_em_from_prepared(C.byref(cfg.grid), r.shape[0] - 1, ptr(t), ptr(pp),
                                              ptr(r), ptr(v), ptr(l_), C.byref(self.s)) == 0
            self.counters = {}
        else:
            rays = np.ascontiguousarray(rays, RAY_DTYPE)
            c = Counters()
            assert lib().orc_em_prepare(C.byref(cfg), rays.shape[0], ptr(rays), C.byref(self.s),
                                        C.byref(c)) == 0
            self.counters = c.as_dict()

    def __del__(self):
        try:
            lib().orc_em_free(C.byref(self.s))
        except Exception:
            pass

    def _arr(self, name, ctype, n):
        a = getattr(self.s, name)
        if not a or n == 0:
            return np.zeros(0, np.dtype(ctype))
        return np.ctypeslib.as_array(C.cast(a, C.POINTER(ctype)), shape=(n,)).copy()

    @property
    def n_sel(self):
        return int(self.s.n_sel)

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    @property
    def row_ptr(self):
        return self._arr("row_ptr", C.c_uint64, self.n_sel + 1)

    @property
    def sel(self):
        return self._arr("sel", C.c_uint32, self.n_sel)

    @property
    def trk4(self):
        return self._arr("trk4", C.c_double, 4 * self.n_sel).reshape(-1, 4)

    @property
    def p(self):
        return self._arr("p", C.c_double, self.n_sel)

    @property
    def pscat(self):
        return self._arr("pscat", C.c_double, 3 * self.n_sel).reshape(-1, 3)

    @property
    def vox(self):
        return self._arr("vox", C.c_uint32, self.nnz)

    @property
    def len(self):
        return self._arr("len", C.c_double, self.nnz)

    @property
    def tail(self):
        return self._arr("tail", C.c_double, self.nnz)

    def prepared(self):
        return self.trk4, self.p, self.row_ptr, self.vox, self.len

    def iterate(self, img, cfg=None):
        cfg = cfg or self.cfg
        img = np.ascontiguousarray(img, np.float64)
        out = np.empty(cfg.grid.n_vox)
        ns = C.c_int64(0)
        cost = lib().orc_em_iterate(C.byref(cfg), C.byref(self.s), ptr(img), ptr(out), C.byref(ns))
        return out, cost, int(ns.value)

    def cij(self, img, cfg=None):
        cfg = cfg or self.cfg
        img = np.ascontiguousarray(img, np.float64)
        c = np.empty(max(self.nnz, 1))
        ci = np.empty(max(self.n_sel, 1))
        lib().orc_em_cij(C.byref(cfg), C.byref(self.s), ptr(img), ptr(c), ptr(ci))
        return c[: self.nnz], ci[: self.n_sel]

    def mj(self):
        m = np.empty(self.cfg.grid.n_vox, np.uint32)
        lib().orc_em_mj(C.byref(self.s), ptr(m))
        return m

    def phi(self, lam, cfg=None):
        cfg = cfg or self.cfg
        return lib().orc_map_phi(C.byref(cfg), C.byref(self.s), ptr(np.ascontiguousarray(lam, np.float64)))

    def grad(self, lam, cfg=None):
        cfg = cfg or self.cfg
        g = np.empty(cfg.grid.n_vox)
        lib().orc_map_grad(C.byref(cfg), C.byref(self.s), ptr(np.ascontiguousarray(lam, np.float64)), ptr(g))
        return g

    def surrogate_grad(self, lam, cfg=None):
        cfg = cfg or self.cfg
        g = np.empty(cfg.grid.n_vox)
        lib().orc_map_surrogate_grad(C.byref(cfg), C.byref(self.s),
                                     ptr(np.ascontiguousarray(lam, np.float64)), ptr(g))
        return g


def mlsd_reconstruct(cfg, rays, init=None, map_=False):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    img = np.empty(cfg.grid.n_vox)
    cost = np.full(max(cfg.num_iteration, 1), np.nan)
    c = Counters()
    init = None if init is None else np.ascontiguousarray(init, np.float64).ravel()
    f = lib().orc_map_reconstruct if map_ else lib().orc_mlsd_reconstruct
    assert f(C.byref(cfg), rays.shape[0], ptr(rays), ptr(init), ptr(img), ptr(cost), C.byref(c)) == 0
    d = c.as_dict()
    return img, cost[: d["iterations_done"]], d


def cubic_solve(p3, p2, p1, p0, tol=1e-8):
    roots = np.zeros(3)
    w = C.c_int(0)
    n = lib().orc_cubic_solve(ptr(roots), p3, p2, p1, p0, tol, C.byref(w))
    return roots[:n], int(w.value)


def map_voxel_step(cfg, mj, last, ml):
    out = np.empty(cfg.grid.n_vox)
    nw = C.c_int64(0)
    lib().orc_map_voxel_step(C.byref(cfg), ptr(np.ascontiguousarray(mj, np.uint32)),
                             ptr(np.ascontiguousarray(last, np.float64)),
                             ptr(np.ascontiguousarray(ml, np.float64)), ptr(out), C.byref(nw))
    return out, int(nw.value)


def map_surrogate(cfg, mj, ljn, ljml, lj, w, lm):
    w = np.ascontiguousarray(w, np.float64)
    lm = np.ascontiguousarray(lm, np.float64)
    return lib().orc_map_surrogate(C.byref(cfg), mj, ljn, ljml, lj, ptr(w), ptr(lm), len(w))


def map_neighbours(cfg, img, j):
    w = np.empty(26)
    lm = np.empty(26)
    n = lib().orc_map_neighbours(C.byref(cfg), ptr(np.ascontiguousarray(img, np.float64)), j,
                                 ptr(w), ptr(lm))
    return w[:n], lm[:n]
