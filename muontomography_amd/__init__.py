"""MI355X (gfx950) reconstruction hot path of muon scattering tomography.

Thin Python handle over libmuto.so (HIP kernels + C ABI, include/muto.h).
The library is the product; this module only moves numpy arrays across the
C ABI.  It raises if the library is missing -- there is no fallback.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from . import _abi
from ._abi import (MUTO_FP32C, MUTO_FP64, MUTO_FP64_STREAM, RAY_DTYPE, Counters, GenBlock, GenCfg, Grid,
                   make_grid, mlsd_cfg, poca_cfg, ptr)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmuto.so")

_lib = None


def lib():
    """Load libmuto.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m muontomography_amd.build`")
        _lib = _abi.bind(C.CDLL(LIB_PATH))
    return _lib


MUTO_MAP_RECOMPUTE_SIDDON = 1
MUTO_MAP_DEVICE_BUFFERS = 2
MUTO_MEM_HOST = 0
MUTO_MEM_DEVICE = 1


def _device_ptr(x, n=None, what="buffer"):
    """Device address of a float64 device buffer -- a devmem.DeviceArray, or
    a contiguous torch tensor on the GPU (after its stream has finished with
    it) -- or None for anything else (host arrays go through numpy)."""
    from .devmem import DeviceArray
    if isinstance(x, DeviceArray):
        if x.dtype != np.float64:
            raise MutoError(f"{what}: device buffers must be float64")
        if n is not None and x.n != n:
            raise MutoError(f"{what}: {x.n} values, expected {n}")
        return x.ptr
    if not (hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)):
        return None
    import torch
    if x.dtype != torch.float64 or not x.is_contiguous():
        raise MutoError(f"{what}: device buffers must be contiguous float64")
    if n is not None and x.numel() != n:
        raise MutoError(f"{what}: {x.numel()} values, expected {n}")
    torch.cuda.current_stream(x.device).synchronize()
    return C.c_void_p(x.data_ptr())


def _device_like(x, n):
    """A new float64 device buffer of n values of the same kind as x."""
    from .devmem import DeviceArray
    if isinstance(x, DeviceArray):
        return DeviceArray(n)
    import torch
    return torch.empty(n, dtype=torch.float64, device=x.device)


MUTO_OPT_GRAPH = 1
MUTO_OPT_TIMING = 2
MUTO_OPT_ACC_MODE = 3
MUTO_OPT_BRICK = 4
MUTO_OPT_FP64_KERNEL = 5
MUTO_OPT_MEDIAN_BRACKETS = 6
# the autotuner's built-in candidate list (muto_em_plan_autotune), in order
def brick_code(shape):
    """MUTO_OPT_BRICK / tuner code of (bx, by, bz) or (bx, by, bz, block):
    bx | by << 8 | bz << 16 | log2(block / 256) << 24 (block 256, 512 or 1024)."""
    bx, by, bz = shape[:3]
    block = shape[3] if len(shape) > 3 else 256
    lb = {256: 0, 512: 1, 1024: 2}[int(block)]
    return int(bx) | int(by) << 8 | int(bz) << 16 | lb << 24


def brick_shape(code):
    """(bx, by, bz, block) of a brick code."""
    c = int(code)
    return (c & 255, (c >> 8) & 255, (c >> 16) & 255, 256 << ((c >> 24) & 255))


# the tuner's built-in candidates (muto_em_plan_autotune with no list)
DEFAULT_BRICKS = [brick_code(s) for s in
                  ((16, 16, 16), (16, 8, 32), (8, 16, 32), (8, 8, 64), (16, 16, 32, 512),
                   (32, 16, 16, 512), (16, 32, 16, 512), (16, 16, 64, 1024), (32, 16, 32, 1024),
                   (16, 16, 32))]


class MutoError(RuntimeError):
    pass


_live_contexts = weakref.WeakSet()


@atexit.register
def _close_contexts():
    """Contexts still open at exit are closed (their batches and plans
    first) while the HIP runtime is still up, not in whatever order the
    interpreter's teardown would free them."""
    for ctx in list(_live_contexts):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """One HIP device + stream (muto_ctx)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        self._plans = weakref.WeakSet()
        self._batches = weakref.WeakSet()
        rc = lib().muto_ctx_create(device, C.byref(self._h))
        if rc != 0:
            raise MutoError(f"muto_ctx_create({device}) failed with status {rc}")
        _live_contexts.add(self)

    def close(self):
        """Destroys the context; plans still open on it are closed first
        (a plan must not outlive its context)."""
        if self._h:
            for b in list(self._batches):
                b.close()
            for p in list(self._plans):
                p.close()
            lib().muto_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().muto_last_error(self._h)
            raise MutoError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def kernel_ms(self):
        """GPU time of the kernels of the last Siddon / PoCA call (ms)."""
        t = C.c_double()
        self._check(lib().muto_ctx_kernel_ms(self._h, C.byref(t)), "muto_ctx_kernel_ms")
        return t.value

    # -- Siddon ---------------------------------------------------------
    def siddon_csr(self, grid, p1, p2):
        """Batched Siddon paths; returns (row_ptr, x, y, z, length)."""
        p1 = np.ascontiguousarray(p1, dtype=np.float64).reshape(-1, 3)
        p2 = np.ascontiguousarray(p2, dtype=np.float64).reshape(-1, 3)
        n = p1.shape[0]
        row = np.zeros(n + 1, dtype=np.uint64)
        self._check(lib().muto_siddon_csr(self._h, C.byref(grid), n, ptr(p1), ptr(p2), ptr(row),
                                          None, None, None, None), "muto_siddon_csr(count)")
        nnz = int(row[-1])
        vx = np.empty(max(nnz, 1), np.uint32)
        vy = np.empty_like(vx)
        vz = np.empty_like(vx)
        ln = np.empty(max(nnz, 1), np.float64)
        self._check(lib().muto_siddon_csr(self._h, C.byref(grid), n, ptr(p1), ptr(p2), ptr(row),
                                          ptr(vx), ptr(vy), ptr(vz), ptr(ln)), "muto_siddon_csr")
        return row, vx[:nnz], vy[:nnz], vz[:nnz], ln[:nnz]

    def siddon_digest(self, grid, p1, p2):
        """Per-segment entry counts (u32) and list digests (u64) without
        storing the lists (muto_siddon_digest)."""
        p1 = np.ascontiguousarray(p1, dtype=np.float64).reshape(-1, 3)
        p2 = np.ascontiguousarray(p2, dtype=np.float64).reshape(-1, 3)
        n = p1.shape[0]
        cnt = np.empty(max(n, 1), np.uint32)
        dig = np.empty(max(n, 1), np.uint64)
        self._check(lib().muto_siddon_digest(self._h, C.byref(grid), n, ptr(p1), ptr(p2), ptr(cnt),
                                             ptr(dig)), "muto_siddon_digest")
        return cnt[:n], dig[:n]

    # -- PoCA -----------------------------------------------------------
    def poca_points(self, grid, rays, eps=1e-8, mode=0):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = rays.shape[0]
        P = np.empty((max(n, 1), 3), np.float64)
        v = np.empty((max(n, 1), 3), np.uint32)
        a = np.empty(max(n, 1), np.uint8)
        c = Counters()
        self._check(lib().muto_poca_points(self._h, C.byref(grid), n, ptr(rays), eps, mode,
                                           ptr(P), ptr(v), ptr(a), C.byref(c)), "muto_poca_points")
        return P[:n], v[:n], a[:n].astype(bool), c.as_dict()

    def poca_reconstruct(self, cfg, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        img = np.empty(cfg.grid.n_vox, np.float64)
        c = Counters()
        self._check(lib().muto_poca_reconstruct(self._h, C.byref(cfg), rays.shape[0], ptr(rays),
                                                ptr(img), C.byref(c)), "muto_poca_reconstruct")
        return img, c.as_dict()

    def rays_from_records(self, records, zlayers, n_middle):
        """Tracker records (n, 2 L + 2) float32 or float64 to rays: NaN
        records dropped, each side fitted on the GPU (MutoFile::getMuonData)."""
        rec = np.ascontiguousarray(records)
        if rec.dtype not in (np.float32, np.float64):
            raise TypeError("records must be float32 or float64")
        z = np.ascontiguousarray(zlayers, np.float64)
        n = rec.shape[0]
        assert rec.shape[1] == 2 * z.shape[0] + 2
        out = np.zeros(max(n, 1), RAY_DTYPE)
        k = C.c_size_t()
        self._check(lib().muto_rays_from_records(self._h, n, z.shape[0], int(n_middle), ptr(z), ptr(rec),
                                                 rec.dtype.itemsize, ptr(out), C.byref(k)),
                    "muto_rays_from_records")
        return out[:k.value]

    def pocut_reconstruct(self, cfg, rays, cut_percentile=1.0):
        """MutoPoCut: PoCA with the per-voxel cut of the largest angles."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        img = np.empty(cfg.grid.n_vox, np.float64)
        c = Counters()
        self._check(lib().muto_pocut_reconstruct(self._h, C.byref(cfg), float(cut_percentile),
                                                 rays.shape[0], ptr(rays), ptr(img), C.byref(c)),
                    "muto_pocut_reconstruct")
        return img, c.as_dict()

    def slr_reconstruct(self, cfg, rays):
        """MutoSLR: straight-line reconstruction."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        img = np.empty(cfg.grid.n_vox, np.float64)
        c = Counters()
        self._check(lib().muto_slr_reconstruct(self._h, C.byref(cfg), rays.shape[0], ptr(rays),
                                               ptr(img), C.byref(c)), "muto_slr_reconstruct")
        return img, c.as_dict()

    # -- MLSD / MAP -----------------------------------------------------
    def _recon(self, fn, cfg, rays, init):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        img = np.empty(cfg.grid.n_vox, np.float64)
        cost = np.full(max(cfg.num_iteration, 1), np.nan)
        c = Counters()
        init = None if init is None else np.ascontiguousarray(init, dtype=np.float64).ravel()
        self._check(fn(self._h, C.byref(cfg), rays.shape[0], ptr(rays), ptr(init), ptr(img),
                       ptr(cost), C.byref(c)), fn.__name__)
        d = c.as_dict()
        return img, cost[:d["iterations_done"]], d

    def mlsd_reconstruct(self, cfg, rays, init=None):
        return self._recon(lib().muto_mlsd_reconstruct, cfg, rays, init)

    def map_reconstruct(self, cfg, rays, init=None):
        return self._recon(lib().muto_map_reconstruct, cfg, rays, init)

    def batch(self, plans):
        """A Batch over EmPlans (mean-update configs sharing the measurement
        errors); the plans must outlive it."""
        return Batch(self, plans)

    def em_plan(self, cfg, rays=None, prepared=None, list_budget=None):
        """list_budget: bytes allowed for the track-major path lists (20 per
        entry); above it the plan keeps none and only the MAP objective
        operator runs, walking the paths again."""
        return EmPlan(self, cfg, rays=rays, prepared=prepared, list_budget=list_budget)


class EmPlan:
    """Device-resident EM problem (muto_em_plan)."""

    def __init__(self, ctx, cfg, rays=None, prepared=None, list_budget=None):
        self.ctx = ctx
        self.cfg = cfg
        self._h = C.c_void_p()
        self._saved = weakref.WeakSet()
        if prepared is not None:
            trk4, p, row, vox, ln = prepared
            self._keep = [np.ascontiguousarray(trk4, np.float64),
                          None if p is None else np.ascontiguousarray(p, np.float64),
                          np.ascontiguousarray(row, np.uint64), np.ascontiguousarray(vox, np.uint32),
                          np.ascontiguousarray(ln, np.float64)]
            t, pp, r, v, l_ = self._keep
            ctx._check(lib().muto_em_plan_create_prepared(ctx._h, C.byref(cfg), r.shape[0] - 1,
                                                          ptr(t), ptr(pp), ptr(r), ptr(v), ptr(l_),
                                                          C.byref(self._h)),
                       "muto_em_plan_create_prepared")
        else:
            rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
            budget = (1 << 64) - 1 if list_budget is None else int(list_budget)
            ctx._check(lib().muto_em_plan_create_ex(ctx._h, C.byref(cfg), rays.shape[0], ptr(rays),
                                                    budget, C.byref(self._h)),
                       "muto_em_plan_create_ex")
        ctx._plans.add(self)

    def close(self):
        if self._h:
            for b in list(self.ctx._batches):  # a batch must not outlive its plans
                if any(p is self for p in b.plans):
                    b.close()
            for sv in list(self._saved):
                sv.close()
            lib().muto_em_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_image(self, img):
        """img: numpy (host) or a float64 device buffer (devmem.DeviceArray
        or a GPU torch tensor), copied in device memory."""
        d = _device_ptr(img, self.cfg.grid.n_vox, "set_image")
        if d is not None:
            self.ctx._check(lib().muto_em_plan_set_image_ex(self._h, d, MUTO_MEM_DEVICE), "set_image_ex")
            return
        img = np.ascontiguousarray(img, np.float64).ravel()
        assert img.size == self.cfg.grid.n_vox
        self.ctx._check(lib().muto_em_plan_set_image(self._h, ptr(img)), "set_image")

    def image(self, out=None):
        """The current image: a new numpy array, or written into `out` (a
        float64 device buffer, filled in device memory)."""
        if out is not None:
            d = _device_ptr(out, self.cfg.grid.n_vox, "image")
            if d is None:
                raise MutoError("image(out=...): out must be a float64 device buffer")
            self.ctx._check(lib().muto_em_plan_get_image_ex(self._h, d, MUTO_MEM_DEVICE), "get_image_ex")
            return out
        img = np.empty(self.cfg.grid.n_vox, np.float64)
        self.ctx._check(lib().muto_em_plan_get_image(self._h, ptr(img)), "get_image")
        return img

    def run(self, iters, precision=MUTO_FP64, map_=False):
        cost = np.full(max(iters, 1), np.nan)
        f = lib().muto_map_plan_run if map_ else lib().muto_em_plan_run
        self.ctx._check(f(self._h, iters, precision, ptr(cost)), f.__name__)
        return cost[: self.counters()["iterations_done"]]

    def counters(self):
        c = Counters()
        self.ctx._check(lib().muto_em_plan_counters(self._h, C.byref(c)), "counters")
        return c.as_dict()

    def voxel_counts(self):
        m = np.empty(self.cfg.grid.n_vox, np.uint32)
        self.ctx._check(lib().muto_em_plan_voxel_counts(self._h, ptr(m)), "voxel_counts")
        return m

    def export(self):
        """(row_ptr, vox, len, tail, trk4, p_r) of the prepared problem."""
        c = self.counters()
        nt, nnz = c["n_scat_in_grid"], c["nnz"]
        row = np.empty(nt + 1, np.uint64)
        vox = np.empty(max(nnz, 1), np.uint32)
        ln = np.empty(max(nnz, 1))
        tl = np.empty(max(nnz, 1))
        t4 = np.empty((max(nt, 1), 4))
        pr = np.empty(max(nt, 1))
        self.ctx._check(lib().muto_em_plan_export(self._h, ptr(row), ptr(vox), ptr(ln), ptr(tl),
                                                  ptr(t4), ptr(pr)), "export")
        return row, vox[:nnz], ln[:nnz], tl[:nnz], t4[:nt], pr[:nt]

    def map_forward(self, lam, precision=MUTO_FP64, recompute=False):
        """Phi(lam) of the MAP objective (muto_map_forward); returns a
        MapSaved whose backward(grad_phi) gives grad_phi * dPhi/dlam."""
        flags = MUTO_MAP_RECOMPUTE_SIDDON if recompute else 0
        d = _device_ptr(lam, self.cfg.grid.n_vox, "map_forward")
        if d is not None:  # a device buffer: gradients in device memory too
            flags |= MUTO_MAP_DEVICE_BUFFERS
            lp = d
        else:
            lam = np.ascontiguousarray(lam, np.float64).ravel()
            assert lam.size == self.cfg.grid.n_vox
            lp = ptr(lam)
        phi = C.c_double()
        h = C.c_void_p()
        self.ctx._check(lib().muto_map_forward(self._h, lp, int(precision), flags,
                                               C.byref(phi), C.byref(h)), "muto_map_forward")
        sv = MapSaved(self, h, phi.value, like=lam if d is not None else None)
        self._saved.add(sv)
        return sv

    def map_value_grad_noalloc(self, lam, grad, phi, recompute=False):
        """Phi into phi[0] and dPhi/dlambda into grad, in one call that
        allocates nothing (muto_map_value_grad_noalloc, FP64): lam and grad
        (n_vox values) and phi (1 value) are float64 device buffers."""
        n = self.cfg.grid.n_vox
        dl = _device_ptr(lam, n, "map_value_grad_noalloc lam")
        dg = _device_ptr(grad, n, "map_value_grad_noalloc grad")
        dp = _device_ptr(phi, 1, "map_value_grad_noalloc phi")
        if dl is None or dg is None or dp is None:
            raise MutoError("map_value_grad_noalloc: lam, grad and phi must be device buffers")
        self.ctx._check(lib().muto_map_value_grad_noalloc(
            self._h, dl, MUTO_MAP_RECOMPUTE_SIDDON if recompute else 0, dp, dg),
            "muto_map_value_grad_noalloc")

    def autotune(self, iters=3, shapes=None):
        """FP32C brick (LDS tile) shape autotuner (muto_em_plan_autotune).
        `shapes`: (bx, by, bz) or (bx, by, bz, block) tuples, or None for the
        built-in list.  Returns the chosen (bx, by, bz, block) and
        {(bx, by, bz, block): ms per iteration}.  The plan's image is left as
        it was."""
        arr = None if shapes is None else np.array([brick_code(sh) for sh in shapes], np.uint32)
        n = 0 if arr is None else arr.size
        ms = np.full(max(n, 16), np.nan)
        best = C.c_uint32(0)
        self.ctx._check(lib().muto_em_plan_autotune(self._h, int(iters), n, ptr(arr), ptr(ms),
                                                    C.byref(best)), "muto_em_plan_autotune")
        if arr is None:
            arr = np.array(DEFAULT_BRICKS, np.uint32)
        return brick_shape(best.value), {brick_shape(v): float(ms[k]) for k, v in enumerate(arr)}

    def set_option(self, option, value):
        """MUTO_OPT_GRAPH, MUTO_OPT_TIMING or MUTO_OPT_ACC_MODE (include/muto.h)."""
        self.ctx._check(lib().muto_em_plan_set_option(self._h, int(option), int(value)),
                        "set_option")

    def timing(self):
        t = np.zeros(5, np.float64)
        self.ctx._check(lib().muto_em_plan_timing(self._h, ptr(t)), "timing")
        return {"forward": t[0], "track": t[1], "backward": t[2], "update": t[3], "iteration": t[4]}


class Batch:
    """Many scans' FP32C mean-update iterations in lock step (muto_batch):
    one launch per pass and one captured graph per iteration for all."""

    def __init__(self, ctx, plans):
        self.ctx = ctx
        self.plans = list(plans)
        self._h = C.c_void_p()
        arr = (C.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        ctx._check(lib().muto_batch_create(ctx._h, len(self.plans), arr, C.byref(self._h)),
                   "muto_batch_create")
        ctx._batches.add(self)

    def run(self, iters):
        """Runs `iters` iterations from the plans' images (written back);
        returns the costs, shape (iters, n_scans)."""
        cost = np.full((max(iters, 1), len(self.plans)), np.nan)
        self.ctx._check(lib().muto_batch_run(self._h, int(iters), ptr(cost)), "muto_batch_run")
        return cost[:iters]

    def stats(self):
        ms = C.c_double()
        c = Counters()
        self.ctx._check(lib().muto_batch_stats(self._h, C.byref(ms), C.byref(c)), "muto_batch_stats")
        return ms.value, c.as_dict()

    def close(self):
        if self._h:
            lib().muto_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MapSaved:
    """Saved context of one muto_map_forward call (autograd-Function style):
    phi, and backward(grad_phi) -> grad_phi * dPhi/dlambda (n_vox doubles)."""

    def __init__(self, plan, h, phi, like=None):
        self.plan = plan
        self._h = h
        self.phi = phi
        self.like = like  # the device lambda of a forward on device buffers, else None

    def backward(self, grad_phi=1.0, out=None):
        """grad_phi * dPhi/dlambda: numpy, or (forward on device buffers) a
        device buffer of the same kind (`out`, or a new one)."""
        n = self.plan.cfg.grid.n_vox
        if self.like is not None:
            if out is None:
                out = _device_like(self.like, n)
            d = _device_ptr(out, n, "backward")
            if d is None:
                raise MutoError("backward(out=...): out must be a float64 device buffer")
            self.plan.ctx._check(lib().muto_map_backward(self._h, float(grad_phi), d),
                                 "muto_map_backward")
            return out
        g = np.empty(n, np.float64)
        self.plan.ctx._check(lib().muto_map_backward(self._h, float(grad_phi), ptr(g)),
                             "muto_map_backward")
        return g

    def close(self):
        if self._h:
            lib().muto_map_saved_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -- synthetic muons (host code in libmuto) -------------------------------
def gen_config(grid, seed, blocks=(), min_voxels=0, max_voxels=0, **kw):
    cfg = GenCfg()
    lib().muto_gen_default(C.byref(cfg), C.byref(grid), seed)
    for k, v in kw.items():
        setattr(cfg, k, v)
    arr = (GenBlock * max(len(blocks), 1))(*blocks)
    cfg._blocks_keep = arr
    cfg.blocks = C.cast(arr, C.POINTER(GenBlock))
    cfg.n_blocks = len(blocks)
    cfg.min_voxels = min_voxels
    cfg.max_voxels = max_voxels
    return cfg


def gen_layout(grid, seed, n, min_mm, max_mm, x0_choices):
    x0 = np.ascontiguousarray(x0_choices, np.float64)
    out = (GenBlock * max(n, 1))()
    rc = lib().muto_gen_layout(C.byref(grid), seed, n, min_mm, max_mm, ptr(x0), len(x0), out)
    if rc != 0:
        raise MutoError(f"muto_gen_layout: status {rc}")
    return list(out)[:n]


def generate(cfg, n):
    rays = np.zeros(n, RAY_DTYPE)
    tried = C.c_int64(0)
    rc = lib().muto_generate(C.byref(cfg), n, ptr(rays), C.byref(tried))
    if rc != 0:
        raise MutoError(f"muto_generate: status {rc}")
    return rays, int(tried.value)


def gen_truth(cfg):
    img = np.empty(cfg.grid.n_vox, np.float64)
    lib().muto_gen_truth(C.byref(cfg), ptr(img))
    return img
