"""MI355X (gfx950) KinectFusion core: the per-frame GPU stages of
PoChang007/Kinect_Fusion_SLAM (bilateral filter, vertex/normal maps,
point-to-plane ICP, projective TSDF integration, voxel-traversal raycast)
as hand-written HIP kernels behind the C ABI in include/kinfu_amd.h.

This package is the Python view of that ABI (ctypes).  Every stage runs in
libkinfu_amd.so; importing without the built library fails.
"""
from ._abi import (KF_FRAME_CURRENT, KF_ICP_STATUS_GATED, KF_MEM_DEVICE,  # noqa: F401
                   KF_MEM_HOST,
                   KF_RAYCAST_REFERENCE,
                   KF_RAYCAST_TRILINEAR, KF_TSDF_F16, KF_TSDF_F32, PRESET_REFERENCE, PRESET_ROOM,
                   PRESET_ROOM_5MM_F16, PRESET_ROOM_CPU, PRESET_ROOM_HD_F16, PRESET_SPHERES_128,
                   PRESETS, STAGES, KinfuContext, KinfuError, Params, SynthScene,
                   VolumeDesc, dist_gate_threshold, exported_symbols, gate_thresholds,
                   jpeg_coefficients, jpeg_info, ldlt6, lib, mat3_inverse,
                   mat4_inverse, orbit_pose, preset, preset_scene, read_depth_file, read_ply, render,
                   spatial_kernel, stream, write_ply)

__all__ = ["KinfuContext", "KinfuError", "VolumeDesc", "Params", "SynthScene", "preset",
           "preset_scene", "orbit_pose", "render", "stream", "gate_thresholds", "mat4_inverse",
           "mat3_inverse", "ldlt6", "spatial_kernel", "lib", "exported_symbols", "jpeg_info",
           "jpeg_coefficients"]
