"""ctypes bindings of the two native libraries built in-tree.

libvirs_hip.so  -- include/virs_fdtd.h (GPU time stepping)
libvirs_host.so -- include/virs_host.h (host grid pipeline)

There is no fallback: if a library is missing the import of the function
that needs it raises, naming the build command.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# VIRS_LIB_DIR selects another build of the same libraries (tuning variants)
LIB_DIR = os.environ.get("VIRS_LIB_DIR") or os.path.join(_HERE, "lib")

_hip = None
_host = None

c_u8p = C.POINTER(C.c_uint8)
c_i16p = C.POINTER(C.c_int16)
c_i32p = C.POINTER(C.c_int32)
c_u32p = C.POINTER(C.c_uint32)
c_i64p = C.POINTER(C.c_int64)
c_fp = C.POINTER(C.c_float)
c_dp = C.POINTER(C.c_double)


class VirsDesc(C.Structure):
    """virs_desc of include/virs_fdtd.h."""
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("h", C.c_float), ("c", C.c_float), ("dt", C.c_float),
        ("precision", C.c_int32), ("sponge_width", C.c_int32), ("sponge_sigma", C.c_float),
        ("device", C.c_int32), ("use_graph", C.c_int32), ("dense_blocks", C.c_int32),
        ("fuse_boundary", C.c_int32), ("rounding", C.c_int32), ("vector_width", C.c_int32),
    ]


class VirsTiming(C.Structure):
    """virs_timing of include/virs_fdtd.h."""
    _fields_ = [
        ("total_ms", C.c_double), ("dense_ms", C.c_double), ("boundary_ms", C.c_double),
        ("sampled_steps", C.c_int64), ("graph_launches", C.c_int64), ("graph_steps", C.c_int32),
        ("tail_graph_steps", C.c_int32), ("single_steps", C.c_int64),
    ]


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m virs_amd.build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    return C.CDLL(path)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def hip():
    """The GPU library, with argument types declared."""
    global _hip
    if _hip is not None:
        return _hip
    L = _load("libvirs_hip.so")
    vp = C.c_void_p
    _sig(L, "virs_desc_default", C.c_int, [C.POINTER(VirsDesc)])
    _sig(L, "virs_create", C.c_int, [C.POINTER(VirsDesc), C.POINTER(vp)])
    _sig(L, "virs_destroy", C.c_int, [vp])
    _sig(L, "virs_last_error", C.c_char_p, [vp])
    _sig(L, "virs_get_constants", C.c_int, [vp, c_fp, c_fp, c_fp, c_fp])
    _sig(L, "virs_upload_flags", C.c_int, [vp, c_u8p])
    _sig(L, "virs_upload_normals", C.c_int, [vp, c_u8p])
    _sig(L, "virs_upload_boundary", C.c_int, [vp, c_u32p, c_u8p, C.c_int64])
    _sig(L, "virs_set_materials", C.c_int, [vp, c_dp, C.c_int32])
    _sig(L, "virs_set_hard_source", C.c_int, [vp, c_dp, C.c_int64])
    _sig(L, "virs_set_soft_sources", C.c_int, [vp, c_i64p, c_dp, C.c_int32, C.c_int64])
    _sig(L, "virs_set_receivers", C.c_int, [vp, c_i64p, C.c_int32])
    _sig(L, "virs_set_early_stop", C.c_int, [vp, C.c_int32, C.c_float])
    _sig(L, "virs_start_simulation", C.c_int, [vp, C.c_int64])
    _sig(L, "virs_step", C.c_int, [vp, C.c_int64, c_i32p])
    _sig(L, "virs_steps_done", C.c_int64, [vp])
    _sig(L, "virs_synchronize", C.c_int, [vp])
    _sig(L, "virs_percent_to_target_rms", C.c_float, [vp])
    _sig(L, "virs_read_receivers", C.c_int, [vp, c_fp])
    _sig(L, "virs_read_receivers_f64", C.c_int, [vp, c_dp])
    _sig(L, "virs_read_field", C.c_int, [vp, C.c_int32, c_fp])
    _sig(L, "virs_read_field_f64", C.c_int, [vp, C.c_int32, c_dp])
    _sig(L, "virs_write_field_f64", C.c_int, [vp, C.c_int32, c_dp])
    _sig(L, "virs_field_exponent", C.c_int, [vp, C.c_int32, c_i32p])
    _sig(L, "virs_memory_bytes", C.c_int, [vp, c_i64p])
    _sig(L, "virs_time_steps", C.c_int, [vp, C.c_int64, C.c_int64, C.POINTER(VirsTiming)])
    _sig(L, "virs_forward_record", C.c_int, [vp, C.c_int64, C.c_int32])
    _sig(L, "virs_adjoint", C.c_int, [vp, c_fp, c_dp])
    _sig(L, "virs_plane_request", C.c_int, [vp, C.c_int64, c_i32p])
    _sig(L, "virs_plane_fetch", C.c_int, [vp, C.c_int32, c_fp])
    _sig(L, "virs_plane_read_noalloc", C.c_int, [vp, C.c_int64, c_fp])
    _hip = L
    return L


def host():
    """The host grid-pipeline library, with argument types declared."""
    global _host
    if _host is not None:
        return _host
    L = _load("libvirs_host.so")
    vp = C.c_void_p
    _sig(L, "virs_mesh_load_obj", C.c_int, [C.c_char_p, C.POINTER(vp)])
    _sig(L, "virs_mesh_from_arrays", C.c_int, [c_fp, c_i32p, C.c_int64, C.POINTER(vp)])
    _sig(L, "virs_mesh_num_triangles", C.c_int64, [vp])
    _sig(L, "virs_mesh_copy", C.c_int, [vp, c_fp, c_i32p])
    _sig(L, "virs_mesh_bounds", C.c_int, [vp, c_fp, c_fp])
    _sig(L, "virs_mesh_free", None, [vp])
    _sig(L, "virs_grid_from_bounds", C.c_int, [c_fp, c_fp, C.c_float, C.c_float, c_fp, c_i64p])
    _sig(L, "virs_reference_cell_size", C.c_float, [])
    _sig(L, "virs_voxelize", C.c_int, [vp, c_fp, C.c_float, c_i64p, c_i16p, C.c_int32, c_u8p, c_u8p, C.c_int32])
    _sig(L, "virs_build_shell", C.c_int, [c_u8p, c_u8p, c_i64p, C.c_int32])
    _sig(L, "virs_compute_normals", C.c_int, [c_u8p, c_u8p, c_i64p, C.c_int32])
    _sig(L, "virs_boundary_list", C.c_int, [c_u8p, c_i64p, c_u32p, c_i64p])
    _sig(L, "virs_boundary_materials", C.c_int, [c_u8p, c_u8p, c_i64p, c_u32p, C.c_int64, c_u8p])
    _sig(L, "virs_rbj_highpass", C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, c_fp])
    _sig(L, "virs_rbj_lowpass", C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, c_fp])
    _sig(L, "virs_ir_resample", C.c_int, [c_fp, C.c_int64, C.c_double, C.c_double, c_fp, C.c_int64])
    _sig(L, "virs_ir_audio", C.c_int, [c_fp, C.c_int64, C.c_float, C.c_int32, c_fp, c_i64p])
    _sig(L, "virs_wav_write_f32", C.c_int, [C.c_char_p, c_fp, C.c_int64, C.c_int32])
    _sig(L, "virs_layer_rgb", C.c_int, [c_fp, c_u8p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, c_u8p,
                                        c_i32p, c_i32p])
    _sig(L, "virs_png_write_rgb", C.c_int, [C.c_char_p, c_u8p, C.c_int32, C.c_int32])
    _host = L
    return L


# Exported symbols each header declares (checked by the CPU test suite).
HIP_SYMBOLS = [
    "virs_desc_default", "virs_create", "virs_destroy", "virs_last_error", "virs_get_constants",
    "virs_upload_flags", "virs_upload_normals", "virs_upload_boundary", "virs_set_materials",
    "virs_set_hard_source", "virs_set_soft_sources", "virs_set_receivers", "virs_set_early_stop",
    "virs_start_simulation", "virs_step", "virs_steps_done", "virs_synchronize",
    "virs_percent_to_target_rms", "virs_read_receivers", "virs_read_receivers_f64",
    "virs_read_field", "virs_read_field_f64", "virs_write_field_f64", "virs_field_exponent",
    "virs_memory_bytes", "virs_time_steps", "virs_forward_record", "virs_adjoint",
    "virs_plane_request", "virs_plane_fetch", "virs_plane_read_noalloc",
]
HOST_SYMBOLS = [
    "virs_mesh_load_obj", "virs_mesh_from_arrays", "virs_mesh_num_triangles", "virs_mesh_copy",
    "virs_mesh_bounds", "virs_mesh_free", "virs_grid_from_bounds", "virs_reference_cell_size",
    "virs_voxelize", "virs_build_shell", "virs_compute_normals", "virs_boundary_list",
    "virs_boundary_materials", "virs_rbj_highpass", "virs_rbj_lowpass",
    "virs_ir_resample", "virs_ir_audio", "virs_wav_write_f32", "virs_layer_rgb", "virs_png_write_rgb",
]
