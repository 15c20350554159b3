"""Python binding of libtsf_hip.so (the gfx950 volume-filter library).

This is a thin ctypes layer over the C ABI declared in include/tsf.h; all
computation happens in the HIP kernels.  Loading fails loudly when the
library has not been built -- there is no CPU fallback.

Volumes are numpy arrays of shape (Z, Y, X); a vector field on the device is
a `Field` (four SoA planes x, y, z, w).
"""
from ._capi import (  # noqa: F401
    TSFError, lib, library_path, Dims, FieldDesc, DeviceArray, Field, Stream, Event,
    STORAGE_F32, STORAGE_F16, STORAGE_BF16, STORAGE_SN16, storage_dtype, sn16_encode, sn16_decode,
    blur_gaussian, create_blur_mask, to_float, create_vector_field, gvf_euler, gvf_fmg,
    hessian_eigen, do_eigen_points, tdf_circle, combine, tdf_round_unorm16, candidates, hp_compact,
    candidate_nms, cube_max, gvf_euler_workspace_size, hp_compact_workspace_size, hp_compact_noalloc,
    gvf_fmg_workspace_size, tdf_circle_workspace_size, device_count, exported_symbols,
    blur_adjoint, vector_field_adjoint, gvf_checkpoint_count, gvf_euler_checkpointed, gvf_euler_adjoint,
    gvf_euler_adjoint_workspace_size, gvf_finish_adjoint, tdf_adjoint, tdf_adjoint_workspace_size, mem_info, mse,
    link_centerpoints, link_centerpoints_workspace_size, graph_components, rasterize_tree_edges,
    segment_inverse_gradient, segment_inverse_gradient_workspace_size, segment_spheres, volume_limits, crop_scan_lines, crop_region, to_float_region,
    CROP_THRESHOLD, CROP_LUNG, tdf_spline, tdf_circle_mixed, tdf_circle_mixed_workspace_size,
    gvf_euler_sr, gvf_euler_sr_workspace_size, gvf_euler_sr_supported, gvf_euler_low_memory, gvf_euler_low_memory_workspace_size,
)
