"""ctypes binding of libvcbct.so (the HIP kernels behind the C ABI of include/vcbct.h).

The library is built in-tree (``python -c "import __graft_entry__ as g; g.build()"`` or
``make -C virtualcbct_amd/csrc``).  Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvcbct.so")

c_int, c_double, c_void_p, c_char_p = C.c_int, C.c_double, C.c_void_p, C.c_char_p
P_D = C.POINTER(C.c_double)
P_F = C.POINTER(C.c_float)
P_U16 = C.POINTER(C.c_uint16)
P_I = C.POINTER(C.c_int)

VCBCT_EINVAL, VCBCT_EHIP, VCBCT_ENOMEM, VCBCT_EIO, VCBCT_ESTATE = -1, -2, -3, -4, -5


class VcbctError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vcbct error {code}: {msg}")
        self.code = code


class vcbct_geometry(C.Structure):
    _fields_ = [("recon_size", c_int), ("recon_scale", c_double), ("recon_slices", c_int),
                ("recon_step", c_double), ("proj_height", c_int), ("proj_width", c_int),
                ("n_angles", c_int), ("x_reg", c_double), ("y_reg", c_double), ("z_reg", c_double),
                ("det_pitch_cm", c_double), ("det_pitch_iso_cm", c_double), ("sad_cm", c_double),
                ("det_pixel_num", c_double), ("num_material", c_int)]


class vcbct_proj_config(C.Structure):
    _fields_ = [("spectrum", C.c_char * 256), ("total_proj", c_int), ("proj_width", c_int),
                ("proj_height", c_int), ("flexmap", C.c_char * 256), ("mat_prefix", C.c_char * 256),
                ("zic", c_int), ("outputname", C.c_char * 256), ("x_scale", c_double), ("y_scale", c_double),
                ("z_scale", c_double), ("recon_size", c_int), ("recon_slices", c_int), ("mat_num", c_int),
                ("mat_type", C.c_char * 16), ("xfactor", c_double), ("x_reg", c_double), ("y_reg", c_double),
                ("z_reg", c_double), ("det_pitch_cm", c_double), ("det_pitch_iso_cm", c_double),
                ("sad_cm", c_double), ("det_pixel_num", c_double), ("sdd_iso_cm", c_double)]


class vcbct_fdk_config(C.Structure):
    _fields_ = [("kvmv", C.c_char * 16), ("total_proj", c_int), ("proj_width", c_int), ("proj_height", c_int),
                ("header_bytes", c_int), ("proj_path", C.c_char * 256), ("num_digits", c_int),
                ("ext", C.c_char * 16), ("flexmap", C.c_char * 256), ("recon_size", c_int),
                ("tgt_class", C.c_char * 16), ("recon_offset_cm", c_int), ("recon_step", c_double),
                ("recon_slices", c_int), ("recon_scale", c_double), ("xfactor", c_int),
                ("output", C.c_char * 256), ("det_pitch_iso_cm", c_double)]


# name -> (restype, argtypes); mirrors include/vcbct.h
SIGNATURES = {
    "vcbct_last_error": (c_char_p, []),
    "vcbct_version": (c_char_p, []),
    "vcbct_projector_create": (c_int, [C.POINTER(c_void_p), C.POINTER(vcbct_geometry), P_D, P_D, P_D, P_D, c_int]),
    "vcbct_projector_destroy": (None, [c_void_p]),
    "vcbct_projector_set_volume_u16": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "vcbct_projector_set_volume_f32": (c_int, [c_void_p, c_void_p, c_void_p]),
    "vcbct_projector_set_spectrum": (c_int, [c_void_p, P_F, P_F, c_int]),
    "vcbct_projector_forward": (c_int, [c_void_p, c_void_p, c_double, c_int, c_int, c_void_p]),
    "vcbct_projector_line_integrals": (c_int, [c_void_p, c_void_p, c_void_p]),
    "vcbct_projector_segment_counts": (c_int, [c_void_p, c_void_p, c_void_p]),
    "vcbct_projector_set_tile": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "vcbct_projector_forward_channels": (c_int, [c_void_p]),
    "vcbct_projector_get_tile": (c_int, [c_void_p, C.POINTER(c_int)]),
    "vcbct_projector_set_adjoint_tile": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "vcbct_reprojection": (c_int, [c_void_p, P_U16, P_F, P_F, c_int, c_double, P_F]),
    "vcbct_reprojection_angles": (c_int, [c_void_p, P_U16, P_F, P_F, c_int, c_double, c_int, P_F]),
    "vcbct_material_table": (c_int, [c_char_p, c_char_p, c_int, C.c_uint, P_U16, P_F, c_int]),
    "vcbct_labels_to_maps": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "vcbct_projector_set_volume_labels": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int,
                                                  c_void_p, c_void_p]),
    "vcbct_projector_get_volume_u16": (c_int, [c_void_p, c_void_p, c_void_p]),
    "vcbct_projector_lengthproj": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "vcbct_ir_update": (c_int, [c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double,
                                c_void_p, c_void_p]),
    "vcbct_fdk_size_class": (c_int, [c_int, P_D, C.c_char_p]),
    "vcbct_fdk_angle_mask": (c_int, [c_int, P_D, c_int, c_double, c_char_p, c_void_p, c_void_p]),
    "vcbct_fdk_half_fan_width": (c_int, [c_int, c_int, P_D]),
    "vcbct_fdk_offset_weight": (c_int, [c_int, c_int, c_int, c_void_p, P_D, P_D, c_double, c_int, c_void_p, P_D, P_D,
                                        c_void_p]),
    "vcbct_detector_counts": (c_int, [C.c_size_t, c_void_p, c_void_p, c_double, C.c_uint64, c_int, c_void_p,
                                      c_void_p]),
    "vcbct_projector_logproj": (c_int, [c_void_p, c_void_p, c_double, c_void_p]),
    "vcbct_projector_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "vcbct_projector_logproj_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "vcbct_projector_logproj_save": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "vcbct_projector_adjoint_channels": (c_int, [c_void_p]),
    "vcbct_projector_set_spectra": (c_int, [c_void_p, c_int, P_F, P_F, P_I]),
    "vcbct_projector_multi_channels": (c_int, [c_void_p]),
    "vcbct_projector_multi_loss_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "vcbct_adam_step": (c_int, [C.c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double,
                                c_double, c_double, c_double, c_void_p]),
    "vcbct_projector_logproj_backward_saved": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "vcbct_prepare_const_map": (c_int, [c_int, c_double, P_D]),
    "vcbct_filter_fft_length": (c_int, [c_int]),
    "vcbct_filter_sinogram": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, P_D, P_D, c_double, c_void_p]),
    "vcbct_reconstruct_image": (c_int, [c_int, c_int, c_int, c_double, c_double, c_void_p, c_int, c_int, c_int,
                                        c_void_p, c_int, P_D, P_D, c_double, c_void_p]),
    "vcbct_reconstruct_image_noalloc": (c_int, [c_int, c_int, c_int, c_double, c_double, c_void_p, c_int, c_int,
                                                c_int, c_void_p, c_int, c_void_p, c_void_p, c_double, c_void_p]),
    "vcbct_filter_sinogram_host": (c_int, [c_int, c_int, c_int, P_F, P_F, P_D, P_D, c_double]),
    "vcbct_reconstruct_image_host": (c_int, [c_int, c_int, c_int, c_double, c_double, P_D, c_int, c_int, c_int, P_F,
                                             c_int, P_D, P_D, c_double]),
    "vcbct_fdk_plan_create": (c_int, [C.POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int, c_int, c_double,
                                      c_double, c_int, P_D, P_D, P_D, c_double, c_int]),
    "vcbct_fdk_plan_destroy": (None, [c_void_p]),
    "vcbct_fdk_plan_filter": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "vcbct_fdk_plan_backproject": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "vcbct_fdk_plan_execute": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "vcbct_fdk_plan_slices_per_thread": (c_int, [c_void_p]),
    "vcbct_fdk_plan_create_half_fan": (c_int, [C.POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int, c_int, c_double,
                                               c_double, c_int, P_D, P_D, P_D, c_double, c_double, c_int]),
    "vcbct_fdk_plan_width": (c_int, [c_void_p]),
    "vcbct_fdk_plan_set_window": (c_int, [c_void_p, c_double]),
    "vcbct_fdk_plan_bp_staged": (c_int, [c_void_p]),
    "vcbct_fdk_offset_weight_host": (c_int, [c_int, c_int, c_int, c_void_p, P_D, P_D, c_double, c_int, c_void_p, P_D,
                                             P_D]),
    "vcbct_detector_counts_host": (c_int, [C.c_size_t, c_void_p, c_void_p, c_double, C.c_uint64, c_int, c_void_p]),
    "vcbct_batch_create": (c_int, [C.POINTER(c_void_p), c_int]),
    "vcbct_batch_destroy": (None, [c_void_p]),
    "vcbct_batch_add_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_double]),
    "vcbct_batch_add_fdk": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.c_longlong]),
    "vcbct_batch_run": (c_int, [c_void_p, c_void_p]),
    "vcbct_batch_capture_graph": (c_int, [c_void_p]),
    "vcbct_batch_launch_graph": (c_int, [c_void_p, c_void_p]),
    "vcbct_batch_blocks": (c_int, [c_void_p]),
    "vcbct_batch_set_grid": (c_int, [c_void_p, c_int]),
    "vcbct_batch_set_tile": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "vcbct_batch_get_tile": (c_int, [c_void_p, C.POINTER(c_int)]),
    "vcbct_autotune_log": (c_int, [C.c_char_p, c_int]),
    "vcbct_read_proj_config": (c_int, [c_char_p, C.POINTER(vcbct_proj_config)]),
    "vcbct_read_flexmap": (c_int, [c_char_p, P_D, P_D, P_D, c_int]),
    "vcbct_proj_slice_window": (c_int, [C.POINTER(vcbct_proj_config), P_I, P_I]),
    "vcbct_load_material_slices": (c_int, [c_char_p, c_int, c_int, c_int, c_int, c_int, P_U16, P_I]),
    "vcbct_read_fdk_config": (c_int, [c_char_p, C.POINTER(vcbct_fdk_config)]),
    "vcbct_read_fdk_angles": (c_int, [c_char_p, c_char_p, c_int, P_D, P_D, P_D]),
    "vcbct_attenuation_table": (c_int, [c_char_p, c_char_p, c_int, c_int, P_F, P_F, c_int]),
    "vcbct_trig_table": (c_int, [c_int, c_int, c_int, P_D, P_D, P_D, c_double, c_double, P_D, P_D, P_D, P_D]),
}

_lib = None


def lib():
    """Loads libvcbct.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libvcbct.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().vcbct_last_error().decode(errors="replace")
        raise VcbctError(rc, msg)
    return rc


def stream_arg(stream=None):
    """The hipStream_t argument of a C-ABI call: `stream` (an integer handle) when given, else torch's current
    stream on the current device."""
    if stream is not None:
        return C.c_void_p(stream)
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(a, t):
    return a.ctypes.data_as(t)
