"""ctypes declarations of the sphx C ABI (include/sphx.h).

`SphxLibrary(path)` binds any shared library exporting the sphx symbols.  The
package itself only ever binds its own HIP build (see `sphinxsys_amd.load`).
"""
import ctypes as C
import os

import numpy as np

SPHX_OK = 0
FP64, MIXED, FP32 = 0, 1, 2
BODY_FLUID, BODY_SOLID, BODY_OBSERVER = 0, 1, 2
RIEMANN_ACOUSTIC, RIEMANN_NONE, RIEMANN_DISSIPATIVE = 0, 1, 2
FLOW_INTERNAL, FLOW_FREE_SURFACE = 0, 1
CORRECTION_NONE, CORRECTION_LINEAR = 0, 1
LIMITER_NONE, LIMITER_TRUNCATED_LINEAR = 0, 1

FIELDS = {
    "Position": (0, "vec"), "VolumetricMeasure": (1, "real"), "Mass": (2, "real"),
    "Density": (3, "real"), "Pressure": (4, "real"), "Velocity": (5, "vec"),
    "Force": (6, "vec"), "ForcePrior": (7, "vec"), "Displacement": (8, "vec"),
    "Compression": (9, "real"), "CompressionRate": (10, "real"),
    "VolumetricMeasureRef": (11, "real"), "CompressionSummation": (12, "real"),
    "NormalDirection": (13, "vec"), "WallVelocity": (14, "vec"),
    "WallAcceleration": (15, "vec"), "OriginalID": (16, "uint"), "SortedID": (17, "uint"),
    "SortSequence": (18, "uint"), "Interpolated": (19, "real"),
    "LinearCorrectionMatrix": (20, "mat"), "Indicator": (21, "int"),
    "PreviousSurfaceIndicator": (22, "int"), "PositionDivergence": (23, "real"),
    "ViscousForce": (24, "vec"), "PreviousViscousForce": (25, "vec"), "KernelGradientIntegral": (26, "vec"),
}

# every symbol include/sphx.h declares
SYMBOLS = [
    "sphx_backend_name", "sphx_last_error", "sphx_context_create", "sphx_context_destroy",
    "sphx_context_synchronize", "sphx_body_create", "sphx_body_destroy", "sphx_body_write",
    "sphx_body_read", "sphx_body_smoothing_length", "sphx_cell_linked_list_create",
    "sphx_cell_linked_list_destroy", "sphx_cell_linked_list_mesh",
    "sphx_update_cell_linked_list", "sphx_update_cell_linked_list_noalloc", "sphx_cell_linked_list_read",
    "sphx_relation_create_inner", "sphx_relation_create_contact", "sphx_relation_destroy",
    "sphx_update_relation", "sphx_relation_read", "sphx_particle_sort", "sphx_gravity_force",
    "sphx_density_regularization", "sphx_advection_step_setup",
    "sphx_update_particle_position", "sphx_compression_summation",
    "sphx_acoustic_step_1st_half", "sphx_acoustic_step_2nd_half", "sphx_acoustic_step_2nd_half_traversal",
    "sphx_acoustic_time_step",
    "sphx_advection_time_step", "sphx_total_mechanical_energy", "sphx_upper_front_in_axis",
    "sphx_interpolate_contact", "sphx_advection_step", "sphx_profile_enable",
    "sphx_profile_read", "sphx_linear_correction_matrix",
    "sphx_acoustic_step_1st_half_corrected", "sphx_acoustic_step_2nd_half_corrected",
    "sphx_free_surface_indication", "sphx_viscous_force", "sphx_kernel_gradient_integral",
    "sphx_transport_velocity_correction",
    "sphx_batch_create", "sphx_batch_destroy", "sphx_batch_set_options", "sphx_batch_add_scene",
    "sphx_batch_remove_scene",
    "sphx_batch_advection_step", "sphx_batch_tuning",
]


class BodyDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("precision", C.c_int), ("kind", C.c_int),
                ("n_particles", C.c_uint32), ("spacing_ref", C.c_double),
                ("h_spacing_ratio", C.c_double), ("rho0", C.c_double), ("c0", C.c_double)]


class AdvectionConfig(C.Structure):
    _fields_ = [("u_ref", C.c_double), ("advection_cfl", C.c_double),
                ("acoustic_cfl", C.c_double), ("riemann", C.c_int), ("flow_type", C.c_int),
                ("sort_now", C.c_int), ("correction", C.c_int), ("correction_alpha", C.c_double),
                ("surface_indication", C.c_int)]


class BatchOptions(C.Structure):
    _fields_ = [("use_graph", C.c_int), ("autotune", C.c_int), ("tile_waves", C.c_int),
                ("stage_min", C.c_uint32), ("substep_margin", C.c_uint32)]


class SphxError(RuntimeError):
    pass


_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint32)


class SphxLibrary:
    """Thin typed binding of one library that exports the sphx C ABI."""

    def __init__(self, path):
        self.path = os.path.abspath(path)
        self.lib = C.CDLL(self.path)
        L = self.lib
        sig = {
            "sphx_backend_name": (C.c_char_p, []),
            "sphx_last_error": (C.c_char_p, []),
            "sphx_context_create": (C.c_int, [C.c_int, _pp]),
            "sphx_context_destroy": (C.c_int, [_vp]),
            "sphx_context_synchronize": (C.c_int, [_vp]),
            "sphx_body_create": (C.c_int, [_vp, C.POINTER(BodyDesc), _pp]),
            "sphx_body_destroy": (C.c_int, [_vp]),
            "sphx_body_write": (C.c_int, [_vp, C.c_int, _vp, C.c_size_t]),
            "sphx_body_read": (C.c_int, [_vp, C.c_int, _vp, C.c_size_t]),
            "sphx_body_smoothing_length": (C.c_int, [_vp, _dp]),
            "sphx_cell_linked_list_create": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_uint32, _pp]),
            "sphx_cell_linked_list_destroy": (C.c_int, [_vp]),
            "sphx_cell_linked_list_mesh": (C.c_int, [_vp, C.POINTER(C.c_int), _dp, _dp]),
            "sphx_update_cell_linked_list": (C.c_int, [_vp]),
            "sphx_update_cell_linked_list_noalloc": (C.c_int, [_vp]),
            "sphx_cell_linked_list_read": (C.c_int, [_vp, _up, C.c_size_t, _up, C.c_size_t]),
            "sphx_relation_create_inner": (C.c_int, [_vp, _vp, _pp]),
            "sphx_relation_create_contact": (C.c_int, [_vp, _vp, _vp, _pp]),
            "sphx_relation_destroy": (C.c_int, [_vp]),
            "sphx_update_relation": (C.c_int, [_vp]),
            "sphx_relation_read": (C.c_int, [_vp, _up, _up, _up, C.c_size_t, C.POINTER(C.c_size_t)]),
            "sphx_particle_sort": (C.c_int, [_vp, _vp]),
            "sphx_gravity_force": (C.c_int, [_vp, _dp]),
            "sphx_density_regularization": (C.c_int, [_vp, C.c_int]),
            "sphx_advection_step_setup": (C.c_int, [_vp]),
            "sphx_update_particle_position": (C.c_int, [_vp]),
            "sphx_compression_summation": (C.c_int, [_vp, _pp, C.c_int]),
            "sphx_acoustic_step_1st_half": (C.c_int, [_vp, _pp, C.c_int, C.c_int, C.c_double]),
            "sphx_acoustic_step_2nd_half": (C.c_int, [_vp, _pp, C.c_int, C.c_int, C.c_double]),
            "sphx_acoustic_step_2nd_half_traversal": (C.c_int, [_vp, _pp, C.c_int, C.c_int, C.c_double]),
            "sphx_acoustic_time_step": (C.c_int, [_vp, C.c_double, _dp]),
            "sphx_advection_time_step": (C.c_int, [_vp, C.c_double, C.c_double, _dp]),
            "sphx_total_mechanical_energy": (C.c_int, [_vp, _dp, _dp]),
            "sphx_upper_front_in_axis": (C.c_int, [_vp, C.c_int, _dp]),
            "sphx_interpolate_contact": (C.c_int, [_vp, C.c_int, _dp, C.c_size_t]),
            "sphx_advection_step": (C.c_int, [_vp, _pp, C.c_int, _vp, C.POINTER(AdvectionConfig),
                                              _dp, _up]),
            "sphx_profile_enable": (C.c_int, [_vp, C.c_int]),
            "sphx_profile_read": (C.c_int, [_vp, C.c_char_p, _dp, C.POINTER(C.c_uint64)]),
            "sphx_linear_correction_matrix": (C.c_int, [_vp, _pp, C.c_int, C.c_double]),
            "sphx_acoustic_step_1st_half_corrected": (C.c_int, [_vp, _pp, C.c_int, C.c_int, C.c_double]),
            "sphx_acoustic_step_2nd_half_corrected": (C.c_int, [_vp, _pp, C.c_int, C.c_int, C.c_double]),
            "sphx_free_surface_indication": (C.c_int, [_vp, _pp, C.c_int]),
            "sphx_viscous_force": (C.c_int, [_vp, _pp, C.c_int, C.c_double, C.c_int]),
            "sphx_kernel_gradient_integral": (C.c_int, [_vp, _pp, C.c_int, C.c_int]),
            "sphx_transport_velocity_correction": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
            "sphx_batch_create": (C.c_int, [C.c_int, _pp]),
            "sphx_batch_destroy": (C.c_int, [_vp]),
            "sphx_batch_set_options": (C.c_int, [_vp, C.POINTER(BatchOptions)]),
            "sphx_batch_add_scene": (C.c_int, [_vp, _vp, _pp, C.c_int, _vp, C.POINTER(AdvectionConfig)]),
            "sphx_batch_remove_scene": (C.c_int, [_vp, C.c_uint32]),
            "sphx_batch_advection_step": (C.c_int, [_vp, _dp, _up, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "sphx_batch_tuning": (C.c_int, [_vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    # -- helpers ---------------------------------------------------------
    @property
    def backend(self):
        return self.lib.sphx_backend_name().decode()

    def check(self, status, what=""):
        if status != SPHX_OK:
            msg = self.lib.sphx_last_error().decode()
            raise SphxError(f"{what}: status {status}: {msg}")

    def call(self, name, *args):
        self.check(getattr(self.lib, name)(*args), name)

    @staticmethod
    def ptr_array(handles):
        arr = (C.c_void_p * max(1, len(handles)))(*[h.value if isinstance(h, C.c_void_p) else h
                                                      for h in handles])
        return C.cast(arr, _pp), arr


def as_dp(a):
    return a.ctypes.data_as(_dp)


def as_up(a):
    return a.ctypes.data_as(_up)
