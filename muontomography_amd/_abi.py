"""ctypes mirror of include/muto.h and include/muto_gen.h (data layout only)."""
import ctypes as C

import numpy as np

MUTO_OK = 0
MUTO_FP64 = 0
MUTO_FP32C = 1
MUTO_FP64_STREAM = 2


class Grid(C.Structure):
    """VoxelGrid (include/MutoTypes.h:25-37)."""
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("_pad", C.c_uint32),
                ("x_min", C.c_double), ("y_min", C.c_double), ("z_min", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]

    @property
    def n_vox(self):
        return int(self.nx) * int(self.ny) * int(self.nz)

    @property
    def shape(self):  # numpy shape of an image in x-fastest order
        return (int(self.nz), int(self.ny), int(self.nx))


def make_grid(n, d, mn):
    n = (n, n, n) if np.isscalar(n) else n
    d = (d, d, d) if np.isscalar(d) else d
    mn = (mn, mn, mn) if np.isscalar(mn) else mn
    return Grid(n[0], n[1], n[2], 0, mn[0], mn[1], mn[2], d[0], d[1], d[2])


# RayData (MutoTypes.h:48-55): 14 doubles
RAY_DTYPE = np.dtype([("ein", "<f8"), ("eout", "<f8"), ("pin", "<f8", 3), ("pout", "<f8", 3),
                      ("din", "<f8", 3), ("dout", "<f8", 3)])
assert RAY_DTYPE.itemsize == 112


class PocaCfg(C.Structure):
    _fields_ = [("grid", Grid), ("length_weighted", C.c_int32), ("straight_line", C.c_int32),
                ("has_energy", C.c_int32), ("log_results", C.c_int32),
                ("total_num_rays", C.c_int64), ("angle_unit_mag", C.c_double), ("eps", C.c_double)]


def poca_cfg(grid, **kw):
    """Defaults of MutoPoCA (src/MutoPoCA.cpp:24-47)."""
    c = PocaCfg()
    c.grid = grid
    c.length_weighted = int(kw.get("length_weighted", False))
    c.straight_line = int(kw.get("straight_line", False))
    c.has_energy = int(kw.get("has_energy", False))
    c.log_results = int(kw.get("log_results", False))
    c.total_num_rays = int(kw.get("total_num_rays", 0))
    c.angle_unit_mag = float(kw.get("angle_unit_mag", 1.0))
    c.eps = float(kw.get("eps", 1e-8))
    return c


class MlsdCfg(C.Structure):
    _fields_ = [("grid", Grid),
                ("update_with_mean", C.c_int32), ("straight_line", C.c_int32),
                ("has_energy", C.c_int32), ("early_stop", C.c_int32),
                ("log_results", C.c_int32), ("mean_percentile", C.c_int32),
                ("num_iteration", C.c_int32), ("initialize_method", C.c_int32),
                ("total_num_rays", C.c_int64),
                ("initialize_value", C.c_double), ("error_xy", C.c_double),
                ("error_dz_outer", C.c_double), ("error_dz_inner", C.c_double),
                ("angle_unit_mag", C.c_double), ("eps", C.c_double),
                ("beta", C.c_double), ("p", C.c_double), ("delta", C.c_double),
                ("add_object_prior", C.c_int32), ("all_neighbours", C.c_int32),
                ("precision", C.c_int32), ("_pad", C.c_int32)]


def mlsd_cfg(grid, **kw):
    """Defaults of MutoMLSD / MutoMAP (src/MutoMLSD.cpp:47-84, src/MutoMAP.cpp:12-19)."""
    c = MlsdCfg()
    c.grid = grid
    c.update_with_mean = int(kw.get("update_with_mean", False))
    c.straight_line = int(kw.get("straight_line", False))
    c.has_energy = int(kw.get("has_energy", False))
    c.early_stop = int(kw.get("early_stop", False))
    c.log_results = int(kw.get("log_results", False))
    c.mean_percentile = int(kw.get("mean_percentile", 100))
    c.num_iteration = int(kw.get("num_iteration", 10))
    c.initialize_method = int(kw.get("initialize_method", 0))
    c.total_num_rays = int(kw.get("total_num_rays", 0))
    c.initialize_value = float(kw.get("initialize_value", 1e-7))
    c.error_xy = float(kw.get("error_xy", 1.0))
    c.error_dz_outer = float(kw.get("error_dz_outer", 1000.0))
    c.error_dz_inner = float(kw.get("error_dz_inner", 1000.0))
    c.angle_unit_mag = float(kw.get("angle_unit_mag", 1.0))
    c.eps = float(kw.get("eps", 1e-8))
    c.beta = float(kw.get("beta", 1.0))
    c.p = float(kw.get("p", 1.0))
    c.delta = float(kw.get("delta", 0.1))
    c.add_object_prior = int(kw.get("add_object_prior", False))
    c.all_neighbours = int(kw.get("all_neighbours", False))
    c.precision = int(kw.get("precision", MUTO_FP64))
    return c


class Counters(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "n_total", "n_not_parallel", "n_scat_in_grid", "n_empty_path", "n_out_of_grid",
        "n_singular", "n_cubic_warn", "n_newton_cap", "nnz", "n_segments", "iterations_done",
        "n_median_miss", "n_map_tile_records", "r1", "r2", "r3")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_ if not n.startswith("r")}


class GenBlock(C.Structure):
    _fields_ = [("x0", C.c_uint32), ("y0", C.c_uint32), ("z0", C.c_uint32),
                ("x1", C.c_uint32), ("y1", C.c_uint32), ("z1", C.c_uint32),
                ("x0_mm", C.c_double)]


class GenCfg(C.Structure):
    _fields_ = [("grid", Grid), ("seed", C.c_uint64), ("tracker_gap", C.c_double),
                ("layer_spacing", C.c_double), ("smear_sigma", C.c_double),
                ("p_median", C.c_double), ("p_sigma_ln", C.c_double), ("p_min", C.c_double),
                ("p_max", C.c_double), ("theta_max_deg", C.c_double), ("air_x0", C.c_double),
                ("n_blocks", C.c_int32), ("min_voxels", C.c_int32), ("max_voxels", C.c_int32),
                ("_pad", C.c_int32), ("blocks", C.POINTER(GenBlock))]


P = C.c_void_p
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)

# name -> (restype, argtypes)
PROTOS = {
    "muto_abi_version": (C.c_int, []),
    "muto_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "muto_ctx_destroy": (C.c_int, [P]),
    "muto_last_error": (C.c_char_p, [P]),
    "muto_ctx_kernel_ms": (C.c_int, [P, C.POINTER(C.c_double)]),
    "muto_siddon_csr": (C.c_int, [P, C.POINTER(Grid), C.c_size_t, P, P, P, P, P, P, P]),
    "muto_siddon_digest": (C.c_int, [P, C.POINTER(Grid), C.c_size_t, P, P, P, P]),
    "muto_poca_points": (C.c_int, [P, C.POINTER(Grid), C.c_size_t, P, C.c_double, C.c_int,
                                   P, P, P, C.POINTER(Counters)]),
    "muto_poca_reconstruct": (C.c_int, [P, C.POINTER(PocaCfg), C.c_size_t, P, P,
                                        C.POINTER(Counters)]),
    "muto_rays_from_records": (C.c_int, [P, C.c_size_t, C.c_int, C.c_int, P, P, C.c_int, P,
                                         C.POINTER(C.c_size_t)]),
    "muto_pocut_reconstruct": (C.c_int, [P, C.POINTER(PocaCfg), C.c_double, C.c_size_t, P, P,
                                         C.POINTER(Counters)]),
    "muto_slr_reconstruct": (C.c_int, [P, C.POINTER(PocaCfg), C.c_size_t, P, P,
                                       C.POINTER(Counters)]),
    "muto_mlsd_reconstruct": (C.c_int, [P, C.POINTER(MlsdCfg), C.c_size_t, P, P, P, P,
                                        C.POINTER(Counters)]),
    "muto_map_reconstruct": (C.c_int, [P, C.POINTER(MlsdCfg), C.c_size_t, P, P, P, P,
                                       C.POINTER(Counters)]),
    "muto_em_plan_create": (C.c_int, [P, C.POINTER(MlsdCfg), C.c_size_t, P, C.POINTER(P)]),
    "muto_em_plan_create_prepared": (C.c_int, [P, C.POINTER(MlsdCfg), C.c_size_t, P, P, P, P, P,
                                               C.POINTER(P)]),
    "muto_em_plan_destroy": (C.c_int, [P]),
    "muto_em_plan_set_image": (C.c_int, [P, P]),
    "muto_em_plan_get_image": (C.c_int, [P, P]),
    "muto_em_plan_set_image_ex": (C.c_int, [P, P, C.c_int]),
    "muto_em_plan_get_image_ex": (C.c_int, [P, P, C.c_int]),
    "muto_em_plan_run": (C.c_int, [P, C.c_int, C.c_int, P]),
    "muto_map_plan_run": (C.c_int, [P, C.c_int, C.c_int, P]),
    "muto_em_plan_counters": (C.c_int, [P, C.POINTER(Counters)]),
    "muto_em_plan_voxel_counts": (C.c_int, [P, P]),
    "muto_em_plan_timing": (C.c_int, [P, P]),
    "muto_em_plan_set_option": (C.c_int, [P, C.c_int, C.c_int]),
    "muto_em_plan_autotune": (C.c_int, [P, C.c_int, C.c_int, P, P, C.POINTER(C.c_uint32)]),
    "muto_em_plan_create_ex": (C.c_int, [P, C.POINTER(MlsdCfg), C.c_size_t, P, C.c_uint64,
                                         C.POINTER(P)]),
    "muto_map_forward": (C.c_int, [P, P, C.c_int, C.c_int, DP, C.POINTER(P)]),
    "muto_map_backward": (C.c_int, [P, C.c_double, P]),
    "muto_map_saved_destroy": (C.c_int, [P]),
    "muto_map_value_grad_noalloc": (C.c_int, [P, P, C.c_int, P, P]),
    "muto_batch_create": (C.c_int, [P, C.c_size_t, P, C.POINTER(P)]),
    "muto_batch_run": (C.c_int, [P, C.c_int, P]),
    "muto_batch_stats": (C.c_int, [P, DP, C.POINTER(Counters)]),
    "muto_batch_destroy": (C.c_int, [P]),
    "muto_em_plan_export": (C.c_int, [P, P, P, P, P, P, P]),
    "muto_gen_default": (None, [C.POINTER(GenCfg), C.POINTER(Grid), C.c_uint64]),
    "muto_gen_layout": (C.c_int, [C.POINTER(Grid), C.c_uint64, C.c_int, C.c_double, C.c_double,
                                  P, C.c_int, C.POINTER(GenBlock)]),
    "muto_generate": (C.c_int, [C.POINTER(GenCfg), C.c_int64, P, C.POINTER(C.c_int64)]),
    "muto_gen_truth": (C.c_int, [C.POINTER(GenCfg), P]),
}


def bind(lib):
    for name, (res, args) in PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def ptr(a):
    """Raw pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)
