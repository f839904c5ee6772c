"""ctypes binding of the FoilCFD C ABI (include/foilcfd.h).

The library is the in-tree ``foil_cfd_amd/libfoilcfd.so`` built for gfx950
(``python -c "import __graft_entry__ as g; g.build()"``). There is no CPU
fallback: loading fails loudly when the shared library is absent.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfoilcfd.so")

FC_OK = 0
FC_ERR_INVALID = 1
FC_ERR_ALLOC = 2
FC_ERR_DEVICE = 3
FC_ERR_STATE = 4
FC_ERR_IO = 5

FC_SOLVER_VERSION = 1

FC_STORAGE_F32 = 0
FC_STORAGE_F16 = 1
FC_STORAGE_BF16 = 2
FC_STORAGE_F64 = 3
FC_STORAGE_F16C = 4

FC_MLP_NPARAMS = 1217

STORAGE_NAMES = {"f32": FC_STORAGE_F32, "f16": FC_STORAGE_F16, "bf16": FC_STORAGE_BF16,
                 "f64": FC_STORAGE_F64, "f16c": FC_STORAGE_F16C}
STORAGE_BYTES = {FC_STORAGE_F32: 4, FC_STORAGE_F16: 2, FC_STORAGE_BF16: 2, FC_STORAGE_F64: 8,
                 FC_STORAGE_F16C: 2}


class FcDims(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)]


class FcScaling(C.Structure):
    _fields_ = [("chord_m", C.c_float), ("airspeed_ms", C.c_float), ("nu_m2s", C.c_float),
                ("chord_cells", C.c_int), ("u_lat", C.c_float), ("dx", C.c_float),
                ("dt", C.c_float), ("nu_lat_target", C.c_float), ("nu_lat", C.c_float),
                ("tau", C.c_float), ("tau_clamped", C.c_int), ("re_target", C.c_float),
                ("re_effective", C.c_float)]


class FcLayout(C.Structure):
    _fields_ = [("dims", FcDims), ("chord_cells", C.c_int), ("anchor_x_frac", C.c_float),
                ("anchor_y_frac", C.c_float), ("anchor_x_cells", C.c_float),
                ("anchor_y_cells", C.c_float), ("anchor_absolute", C.c_int)]


class FcPatchBox(C.Structure):
    _fields_ = [("x0", C.c_int), ("y0", C.c_int), ("x1", C.c_int), ("y1", C.c_int)]


class FcRefinementInfo(C.Structure):
    _fields_ = [("active", C.c_int), ("factor", C.c_int), ("box", FcPatchBox),
                ("fine_dims", FcDims), ("fine_scaling", FcScaling), ("device_bytes", C.c_double),
                ("forces_from_fine", C.c_int)]


class FcForceReadout(C.Structure):
    _fields_ = [("cl", C.c_float), ("cd", C.c_float), ("lift_to_drag", C.c_float),
                ("valid", C.c_int), ("flow_throughs", C.c_float),
                ("cl_avg", C.c_float), ("cl_min", C.c_float), ("cl_max", C.c_float),
                ("cl_median", C.c_float), ("cd_avg", C.c_float), ("cd_min", C.c_float),
                ("cd_max", C.c_float), ("cd_median", C.c_float), ("ld_avg", C.c_float),
                ("ld_min", C.c_float), ("ld_max", C.c_float), ("ld_median", C.c_float)]


class FcPerfStats(C.Structure):
    _fields_ = [("last_steps_per_frame", C.c_int), ("last_step_ms", C.c_double),
                ("mlups", C.c_double), ("population_bytes_per_step", C.c_double)]


class FcWallModelReadout(C.Structure):
    _fields_ = [("enabled", C.c_int), ("mode", C.c_int), ("listed", C.c_int), ("cells", C.c_int),
                ("mean_yplus", C.c_float), ("max_yplus", C.c_float), ("clamped_frac", C.c_float),
                ("excluded_vg", C.c_int), ("degenerate", C.c_int), ("mlp_cells", C.c_int),
                ("from_fine", C.c_int)]


class FcMlpParams(C.Structure):
    _fields_ = [("in_mean", C.c_float * 4), ("in_inv_std", C.c_float * 4),
                ("env_lo", C.c_float * 4), ("env_hi", C.c_float * 4),
                ("out_mean", C.c_float), ("out_std", C.c_float),
                ("w1", C.c_uint16 * (32 * 3)), ("b1", C.c_float * 32),
                ("w2", C.c_uint16 * (32 * 32)), ("b2", C.c_float * 32),
                ("w3", C.c_uint16 * 32), ("b3", C.c_float)]


class FcGroupInfo(C.Structure):
    _fields_ = [("members", C.c_int), ("cells", C.c_longlong), ("waves_per_block", C.c_int),
                ("slab_order", C.c_int), ("use_graph", C.c_int), ("cache_budget", C.c_longlong),
                ("sub_batches", C.c_int), ("last_step_ms", C.c_double),
                ("glups", C.c_double)]


class FcVGParams(C.Structure):
    _fields_ = [("type", C.c_int), ("x_c", C.c_float), ("height_c", C.c_float),
                ("length_h", C.c_float), ("beta_deg", C.c_float), ("pitch_c", C.c_float),
                ("gap_h", C.c_float), ("count", C.c_int), ("common_flow_down", C.c_int),
                ("enabled", C.c_int)]


class FcVaneSlab(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("n_axis", C.c_float * 3), ("d_axis", C.c_float * 3),
                ("b_axis", C.c_float * 3), ("half_height", C.c_float),
                ("half_length", C.c_float), ("half_thick", C.c_float)]


class FcSnapshotKey(C.Structure):
    _fields_ = [("geometry_id", C.c_char_p), ("aoa_bucket", C.c_int), ("u_lat", C.c_float),
                ("dims", FcDims), ("solver_version", C.c_int)]


class FcDelta99Sample(C.Structure):
    _fields_ = [("x_c", C.c_float), ("delta99_c", C.c_float), ("ue_edge", C.c_float), ("valid", C.c_int)]


class FcPlaneCirculation(C.Structure):
    _fields_ = [("pos", C.c_float), ("neg", C.c_float)]


class FcBandCirculation(C.Structure):
    _fields_ = [("zone", FcPlaneCirculation), ("ambient", FcPlaneCirculation), ("zone_cols", C.c_int),
                ("ambient_cols", C.c_int)]


class FcVGAuditReadout(C.Structure):
    _fields_ = [("valid", C.c_int), ("gamma_measured", C.c_float), ("gamma_predicted", C.c_float),
                ("ratio", C.c_float), ("plane_xc", C.c_float)]


class FcInitOptions(C.Structure):
    _fields_ = [("storage", C.c_int), ("periodic_x", C.c_int), ("periodic_y", C.c_int)]


_vp = C.c_void_p
_u8p = C.POINTER(C.c_uint8)
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_llp = C.POINTER(C.c_longlong)

# name -> (restype, argtypes); every symbol include/foilcfd.h declares.
SIGNATURES = {
    "fc_version": (C.c_char_p, []),
    "fc_device_count": (C.c_int, []),
    "fc_compute_scaling": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_int, C.c_float,
                                     C.POINTER(FcScaling)]),
    "fc_max_effective_re": (C.c_float, [C.c_int, C.c_float]),
    "fc_refined_scaling": (C.c_int, [C.POINTER(FcScaling), C.c_int, C.POINTER(FcScaling)]),
    "fc_fine_tau_for": (C.c_float, [C.c_float, C.c_int]),
    "fc_ramped_tau": (C.c_float, [C.POINTER(FcScaling), C.c_longlong, C.c_int]),
    "fc_ramped_u": (C.c_float, [C.POINTER(FcScaling), C.c_longlong, C.c_int]),
    "fc_force_to_coefficient": (C.c_float, [C.POINTER(FcScaling), C.c_float, C.c_int]),
    "fc_flow_through_steps": (C.c_float, [C.POINTER(FcScaling), C.c_int]),
    "fc_generate_naca4": (C.c_int, [C.c_char_p, C.c_int, _fp, C.c_int, _ip, C.c_char_p, C.c_int]),
    "fc_build_boundary_flags": (C.c_int, [FcDims, _u8p]),
    "fc_voxelize_airfoil": (C.c_int, [_fp, C.c_int, C.c_float, C.POINTER(FcLayout), _u8p]),
    "fc_close_trailing_edge_gaps": (C.c_int, [FcDims, _u8p]),
    "fc_build_clean_foil_flags": (C.c_int, [_fp, C.c_int, C.c_float, C.POINTER(FcLayout), _u8p]),
    "fc_make_fine_layout": (C.c_int, [C.POINTER(FcLayout), C.POINTER(FcPatchBox), C.c_int,
                                      C.POINTER(FcLayout)]),
    "fc_fine_dims": (C.c_int, [C.POINTER(FcPatchBox), FcDims, C.c_int, C.POINTER(FcDims)]),
    "fc_stamp_interface_shell": (C.c_int, [FcDims, _u8p]),
    "fc_build_fine_patch_flags": (C.c_int, [_fp, C.c_int, C.c_float, C.POINTER(FcLayout),
                                            C.POINTER(FcPatchBox), C.c_int, _u8p]),
    "fc_derive_patch_box": (C.c_int, [FcDims, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(FcPatchBox)]),
    "fc_surface_frame_at": (C.c_int, [_fp, C.c_int, C.c_int, C.c_float, C.c_int, _fp]),
    "fc_build_flags_with_vgs": (C.c_int, [_fp, C.c_int, C.c_int, C.c_float, C.POINTER(FcLayout),
                                          C.POINTER(FcVGParams), C.c_int, _u8p, _u8p,
                                          C.POINTER(FcVaneSlab), C.c_int, _ip]),
    "fc_build_vane_qlinks": (C.c_int, [FcDims, _u8p, _u8p, C.POINTER(FcVaneSlab), C.c_int, _u8p,
                                       C.POINTER(C.c_uint32)]),
    "fc_ray_obb_entry": (C.c_float, [_fp, _fp, C.POINTER(FcVaneSlab)]),
    "fc_create": (_vp, []),
    "fc_destroy": (None, [_vp]),
    "fc_init": (C.c_int, [_vp, FcDims, C.POINTER(FcScaling), _u8p, C.POINTER(FcInitOptions),
                          C.c_char_p, C.c_int]),
    "fc_shutdown": (None, [_vp]),
    "fc_reset": (C.c_int, [_vp]),
    "fc_set_flags": (C.c_int, [_vp, _u8p]),
    "fc_set_startup_ramp": (C.c_int, [_vp, C.c_int]),
    "fc_set_force_ema_window": (C.c_int, [_vp, C.c_float]),
    "fc_set_smagorinsky_cs": (C.c_int, [_vp, C.c_float]),
    "fc_set_step_variant": (C.c_int, [_vp, C.c_int]),
    "fc_set_stochastic_rounding": (C.c_int, [_vp, C.c_int]),
    "fc_get_stochastic_rounding": (C.c_int, [_vp]),
    "fc_last_error": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "fc_init_refinement": (C.c_int, [_vp, C.POINTER(FcPatchBox), C.c_int, _u8p, C.c_char_p,
                                     C.c_int]),
    "fc_shutdown_refinement": (C.c_int, [_vp]),
    "fc_set_refined_flags": (C.c_int, [_vp, _u8p]),
    "fc_get_refinement_info": (C.c_int, [_vp, C.POINTER(FcRefinementInfo)]),
    "fc_init_finer_refinement": (C.c_int, [_vp, C.POINTER(FcPatchBox), C.c_int, _u8p, C.c_char_p,
                                           C.c_int]),
    "fc_shutdown_finer_refinement": (C.c_int, [_vp]),
    "fc_get_finer_refinement_info": (C.c_int, [_vp, C.POINTER(FcRefinementInfo)]),
    "fc_download_finer_f": (C.c_int, [_vp, _vp, C.c_int]),
    "fc_upload_finer_f": (C.c_int, [_vp, _vp, C.c_int]),
    "fc_download_fine_f": (C.c_int, [_vp, _vp, C.c_int]),
    "fc_upload_fine_f": (C.c_int, [_vp, _vp, C.c_int]),
    "fc_step": (C.c_int, [_vp, C.c_int]),
    "fc_synchronize": (C.c_int, [_vp]),
    "fc_forces": (C.c_int, [_vp, C.POINTER(FcForceReadout)]),
    "fc_force_raw": (C.c_int, [_vp, _dp, _llp]),
    "fc_force_now": (C.c_int, [_vp, _dp]),
    "fc_force_now_noalloc": (C.c_int, [_vp, _dp]),
    "fc_force_fold_create": (_vp, []),
    "fc_force_fold_destroy": (None, [_vp]),
    "fc_force_fold_clear": (C.c_int, [_vp]),
    "fc_force_fold_add": (C.c_int, [_vp, C.c_longlong, C.c_float, C.c_float, C.c_double, C.c_double]),
    "fc_force_fold_readout": (C.c_int, [_vp, C.POINTER(FcScaling), C.c_int, C.c_longlong, C.c_int,
                                        C.POINTER(FcForceReadout)]),
    "fc_mass": (C.c_int, [_vp, _dp]),
    "fc_read_macro": (C.c_int, [_vp, _fp, _fp, _fp, _fp]),
    "fc_read_flags": (C.c_int, [_vp, _u8p]),
    "fc_download_f": (C.c_int, [_vp, _vp, C.c_int]),
    "fc_upload_f": (C.c_int, [_vp, _vp, C.c_int]),
    "fc_notify_snapshot_restored": (C.c_int, [_vp, C.c_int]),
    "fc_nan_detected": (C.c_int, [_vp]),
    "fc_nan_kind": (C.c_int, [_vp]),
    "fc_get_perf_stats": (C.c_int, [_vp, C.POINTER(FcPerfStats)]),
    "fc_step_count": (C.c_longlong, [_vp]),
    "fc_current_tau": (C.c_float, [_vp]),
    "fc_flow_throughs_completed": (C.c_float, [_vp]),
    "fc_get_dims": (C.c_int, [_vp, C.POINTER(FcDims)]),
    "fc_get_scaling": (C.c_int, [_vp, C.POINTER(FcScaling)]),
    "fc_storage": (C.c_int, [_vp]),
    "fc_device_fields": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                   C.POINTER(_vp), C.POINTER(_vp)]),
    "fc_device_bytes": (C.c_longlong, [_vp]),
    "fc_build_wall_cell_list": (C.c_int, [FcDims, _u8p, _u8p, C.c_int, _llp, _llp, _fp, _fp, _fp,
                                          _fp, C.POINTER(C.c_uint32), C.c_int, _ip]),
    "fc_set_wall_model": (C.c_int, [_vp, C.c_int]),
    "fc_set_surface_reference": (C.c_int, [_vp, _u8p]),
    "fc_get_wall_model_readout": (C.c_int, [_vp, C.POINTER(FcWallModelReadout)]),
    "fc_read_wall_slip": (C.c_int, [_vp, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16),
                                    C.POINTER(C.c_uint16), _fp]),
    "fc_wall_update_now": (C.c_int, [_vp]),
    "fc_wall_cell_count": (C.c_int, [_vp]),
    "fc_set_wall_harvest": (C.c_int, [_vp, C.c_int]),
    "fc_wall_harvest": (C.c_int, [_vp, _fp, C.c_int]),
    "fc_set_wall_mlp": (C.c_int, [_vp, C.POINTER(FcMlpParams)]),
    "fc_mlp_trainer_create": (_vp, [_fp, C.c_longlong, C.c_uint32, C.c_char_p, C.c_int]),
    "fc_mlp_trainer_destroy": (None, [_vp]),
    "fc_mlp_trainer_samples": (C.c_longlong, [_vp]),
    "fc_mlp_trainer_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, _fp]),
    "fc_mlp_trainer_gradient": (C.c_int, [_vp, _fp, _ip, C.c_int, _fp, _fp]),
    "fc_mlp_trainer_get_params": (C.c_int, [_vp, _fp]),
    "fc_mlp_trainer_set_params": (C.c_int, [_vp, _fp]),
    "fc_mlp_trainer_data": (C.c_int, [_vp, _fp, _fp]),
    "fc_mlp_trainer_export": (C.c_int, [_vp, C.POINTER(FcMlpParams)]),
    "fc_mlp_infer": (C.c_int, [C.POINTER(FcMlpParams), _fp, C.c_int, _fp]),
    "fc_group_create": (_vp, [C.POINTER(_vp), C.c_int, C.c_char_p, C.c_int]),
    "fc_group_destroy": (None, [_vp]),
    "fc_group_step": (C.c_int, [_vp, C.c_int]),
    "fc_group_set_launch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "fc_group_set_cache_budget": (C.c_int, [_vp, C.c_longlong]),
    "fc_group_last_step_ms": (C.c_int, [_vp, _dp]),
    "fc_group_get_info": (C.c_int, [_vp, C.POINTER(FcGroupInfo)]),
    "fc_group_last_error": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "fc_group_autotune": (C.c_int, [_vp, C.c_int, C.POINTER(FcGroupInfo)]),
    "fc_set_qlibb_enabled": (C.c_int, [_vp, C.c_int]),
    "fc_set_qlinks": (C.c_int, [_vp, _u8p, C.POINTER(C.c_uint32), C.c_int, C.c_int]),
    "fc_set_fine_qlinks": (C.c_int, [_vp, _u8p, C.POINTER(C.c_uint32), C.c_int, C.c_int]),
    "fc_set_finer_qlinks": (C.c_int, [_vp, _u8p, C.POINTER(C.c_uint32), C.c_int, C.c_int]),
    "fc_qlibb_readout": (C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "fc_particles_seed": (C.c_int, [_vp, C.c_int, C.c_uint]),
    "fc_particles_advect": (C.c_int, [_vp, C.c_float, C.c_uint]),
    "fc_advect_particles": (C.c_int, [_vp, _vp, C.c_int, C.c_float, C.c_uint]),
    "fc_particles_read": (C.c_int, [_vp, _fp]),
    "fc_particles_write": (C.c_int, [_vp, _fp]),
    "fc_particles_count": (C.c_int, [_vp]),
    "fc_particles_set_params": (C.c_int, [_vp, C.c_float, C.c_float]),
    "fc_apply_edited_flags": (C.c_int, [_vp, _u8p]),
    "fc_last_edit_count": (C.c_longlong, [_vp]),
    "fc_restore_from_macroscopic": (C.c_int, [_vp, _fp, _fp, _fp, _fp]),
    "fc_snapshot_aoa_bucket": (C.c_int, [C.c_float]),
    "fc_snapshot_key_hash": (C.c_uint64, [C.POINTER(FcSnapshotKey)]),
    "fc_snapshot_key_string": (C.c_int, [C.POINTER(FcSnapshotKey), C.c_char_p, C.c_int]),
    "fc_snapshot_key_file_name": (C.c_int, [C.POINTER(FcSnapshotKey), C.c_char_p, C.c_int]),
    "fc_snapshot_key_equal": (C.c_int, [C.POINTER(FcSnapshotKey), C.POINTER(FcSnapshotKey)]),
    "fc_hash_flags": (C.c_uint64, [_u8p, C.c_uint64]),
    "fc_vram_snapshot_create": (_vp, []),
    "fc_vram_snapshot_destroy": (None, [_vp]),
    "fc_vram_snapshot_capture": (C.c_int, [_vp, _vp, C.POINTER(FcSnapshotKey), C.c_char_p, C.c_int]),
    "fc_vram_snapshot_restore": (C.c_int, [_vp, _vp]),
    "fc_vram_snapshot_has_data": (C.c_int, [_vp]),
    "fc_vram_snapshot_get_key": (C.c_int, [_vp, C.POINTER(FcSnapshotKey), C.c_char_p, C.c_int]),
    "fc_vram_snapshot_flags_hash": (C.c_uint64, [_vp]),
    "fc_vram_snapshot_bytes": (C.c_longlong, [_vp]),
    "fc_vram_snapshot_release": (None, [_vp]),
    "fc_compact_snapshot_create": (_vp, []),
    "fc_compact_snapshot_destroy": (None, [_vp]),
    "fc_compact_snapshot_capture": (C.c_int, [_vp, _vp, C.POINTER(FcSnapshotKey), C.c_char_p, C.c_int]),
    "fc_compact_snapshot_set": (C.c_int, [_vp, C.POINTER(FcSnapshotKey), C.c_uint64, C.POINTER(FcScaling),
                                          _fp, _fp, _fp, _fp, C.c_longlong]),
    "fc_compact_snapshot_restore": (C.c_int, [_vp, _vp]),
    "fc_compact_snapshot_save": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_int]),
    "fc_compact_snapshot_load": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_int]),
    "fc_compact_snapshot_has_data": (C.c_int, [_vp]),
    "fc_compact_snapshot_get_key": (C.c_int, [_vp, C.POINTER(FcSnapshotKey), C.c_char_p, C.c_int]),
    "fc_compact_snapshot_flags_hash": (C.c_uint64, [_vp]),
    "fc_compact_snapshot_scaling": (C.c_int, [_vp, C.POINTER(FcScaling)]),
    "fc_compact_snapshot_fields": (C.c_int, [_vp, _fp, _fp, _fp, _fp]),
    "fc_snapshot_cache_open": (_vp, [C.c_char_p, C.c_uint64]),
    "fc_snapshot_cache_close": (None, [_vp]),
    "fc_snapshot_cache_find": (C.c_int, [_vp, C.POINTER(FcSnapshotKey), C.c_char_p, C.c_int]),
    "fc_snapshot_cache_store": (C.c_int, [_vp, _vp, C.c_char_p, C.c_int]),
    "fc_snapshot_cache_store_async": (C.c_int, [_vp, _vp]),
    "fc_snapshot_cache_flush": (C.c_longlong, [_vp]),
    "fc_snapshot_cache_set_budget": (C.c_int, [_vp, C.c_uint64]),
    "fc_snapshot_cache_used_bytes": (C.c_uint64, [_vp]),
    "fc_wall_cell_count_level": (C.c_int, [_vp, C.c_int]),
    "fc_read_wall_slip_level": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16),
                                          C.POINTER(C.c_uint16), _fp]),
    "fc_set_refined_surface_reference": (C.c_int, [_vp, _u8p]),
    "fc_set_refined_finer_surface_reference": (C.c_int, [_vp, _u8p]),
    "fc_adaptive_steps_for_budget": (C.c_int, [_vp, C.c_double]),
    "fc_seed_from_coarse": (C.c_int, [_vp, _vp]),
    "fc_preconverge": (C.c_int, [_vp, FcDims, C.POINTER(FcScaling), _u8p, C.POINTER(FcInitOptions),
                                 C.POINTER(C.c_longlong), C.c_char_p, C.c_int]),
    "fc_extract_suction_delta99": (C.c_int, [_vp, _fp, C.c_int, C.POINTER(FcDelta99Sample)]),
    "fc_separation_onset_xc": (C.c_int, [_vp, _fp]),
    "fc_download_crossflow_plane": (C.c_int, [_vp, C.c_int, _fp, _fp]),
    "fc_suction_surface_y": (C.c_int, [_vp, C.c_int]),
    "fc_lattice_x_for_chord_station": (C.c_int, [_vp, C.c_float]),
    "fc_band_circulation_live": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(FcBandCirculation)]),
    "fc_wendt_circulation_lattice": (C.c_float, [C.c_float] * 5),
    "fc_integrate_streamwise_circulation": (C.c_int, [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int,
                                                      C.POINTER(FcPlaneCirculation)]),
    "fc_integrate_band_circulation": (C.c_int, [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.POINTER(FcBandCirculation)]),
    "fc_audit_vg_vortex_strength": (C.c_int, [_vp, C.POINTER(FcVGParams), C.c_int, C.c_float, C.c_float,
                                              C.POINTER(FcVGAuditReadout)]),
    "fc_set_kernel_timing": (C.c_int, [_vp, C.c_int]),
    "fc_kernel_timing": (C.c_int, [_vp, _ip, _dp]),
    "fc_kernel_timing_level": (C.c_int, [_vp, C.c_int, _ip, _dp]),
    "fc_time_steps": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "fc_probe_aa_traffic": (C.c_int, [FcDims, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "fc_probe_store_rounding": (C.c_int, [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_int,
                                           C.c_longlong, C.POINTER(C.c_uint16)]),
    "fc_probe_copy": (C.c_int, [C.c_size_t, C.c_int, _dp]),
}

_LIB = None


def load(path: str | None = None) -> C.CDLL:
    """Load libfoilcfd.so and bind every C-ABI symbol; raises if missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"FoilCFD HIP library not built: {p} is missing. Run "
            "`python -c 'import __graft_entry__ as g; g.build()'` first.")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(rc: int, what: str, msg: str = "") -> None:
    if rc != FC_OK:
        raise RuntimeError(f"{what} failed with status {rc}{': ' + msg if msg else ''}")
