"""Python mirror of the SPHinXsys computing-kernel (CK) dynamics used by the
weakly-compressible fluid step, over the sphx C ABI.

Class names, constructor arguments and exec() semantics follow the reference:
  UpdateCellLinkedList  shared_ck/particle_dynamics/configuration_dynamics/update_cell_linked_list.h:39-84
  UpdateRelation        configuration_dynamics/update_body_relation.h:97-213
  ParticleSortCK        configuration_dynamics/particle_sort_ck.h:40-101
  GravityForceCK        general_dynamics/force_prior_ck.h:40-93
  CompressionSummation, DensityRegularization   fluid_dynamics/density_regularization.h:46-184
  AcousticStep1stHalf / AcousticStep2ndHalf     fluid_dynamics/acoustic_step_{1st,2nd}_half.h
  AcousticTimeStepCK, AdvectionTimeStepCK, AdvectionStepSetup, UpdateParticlePosition
                        fluid_dynamics/fluid_time_step_ck.h:42-184
  TotalMechanicalEnergyCK, UpperFrontInAxisDirectionCK  general_dynamics/general_reduce_ck.h:69-222
  Interpolation<Contact<Real>>                  general_dynamics/interpolation_dynamics.hpp:46-60
  LinearCorrectionMatrix                        general_dynamics/kernel_correction_ck.h:40-120
  FreeSurfaceIndicationCK                       general_dynamics/surface_indication/surface_indication_ck.h
  ViscousForceCK                                fluid_dynamics/viscous_force.h:43-127
  KernelGradientIntegral                        general_dynamics/kernel_gradient_integral.h:36-96
  TransportVelocityCorrectionCK                 fluid_dynamics/transport_velocity_correction_ck.h
Every object takes the SphxLibrary it runs on; the product binding is
`sphinxsys_amd.load()` (HIP).  Errors raise SphxError with the library's message.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from ._capi import as_dp, as_up


class Context:
    def __init__(self, lib, device=0):
        self.lib = lib
        self.h = C.c_void_p()
        lib.call("sphx_context_create", device, C.byref(self.h))

    def synchronize(self):
        self.lib.call("sphx_context_synchronize", self.h)

    def profile(self, enable=True, passes_only=False):
        """Start (or stop) HIP-event timing of the library's kernel stages (all of
        them, or only the two acoustic passes)."""
        self.lib.call("sphx_profile_enable", self.h, (2 if passes_only else 1) if enable else 0)

    def profile_read(self, stage):
        """(total milliseconds, launches) of one stage since profile() was called."""
        ms, n = C.c_double(), C.c_uint64()
        self.lib.call("sphx_profile_read", self.h, stage.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if self.h:
            self.lib.lib.sphx_context_destroy(self.h)
            self.h = C.c_void_p()


class Body:
    """A particle set (FluidBody / SolidBody / ObserverBody) and its fields."""

    def __init__(self, ctx, dim, n, spacing_ref, kind=capi.BODY_FLUID, precision=capi.FP64,
                 h_spacing_ratio=1.3, rho0=1.0, c0=1.0):
        self.ctx, self.lib = ctx, ctx.lib
        self.dim, self.n, self.kind, self.precision = dim, int(n), kind, precision
        self.desc = capi.BodyDesc(dim, precision, kind, int(n), float(spacing_ref),
                                  float(h_spacing_ratio), float(rho0), float(c0))
        self.h = C.c_void_p()
        self.lib.call("sphx_body_create", ctx.h, C.byref(self.desc), C.byref(self.h))

    @property
    def smoothing_length(self):
        v = C.c_double()
        self.lib.call("sphx_body_smoothing_length", self.h, C.byref(v))
        return v.value

    def write(self, name, values):
        fid, kind = capi.FIELDS[name]
        if kind == "uint":
            a = np.ascontiguousarray(values, dtype=np.uint32).reshape(self.n)
        elif kind == "int":
            a = np.ascontiguousarray(values, dtype=np.int32).reshape(self.n)
        elif kind == "mat":
            a = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n, self.dim, self.dim)
        elif kind == "vec":
            a = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n, self.dim)
        else:
            a = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n)
        self.lib.call("sphx_body_write", self.h, fid, a.ctypes.data_as(C.c_void_p), self.n)

    def read(self, name):
        fid, kind = capi.FIELDS[name]
        if kind == "uint":
            a = np.empty(self.n, dtype=np.uint32)
        elif kind == "int":
            a = np.empty(self.n, dtype=np.int32)
        elif kind == "mat":
            a = np.empty((self.n, self.dim, self.dim), dtype=np.float64)
        elif kind == "vec":
            a = np.empty((self.n, self.dim), dtype=np.float64)
        else:
            a = np.empty(self.n, dtype=np.float64)
        self.lib.call("sphx_body_read", self.h, fid, a.ctypes.data_as(C.c_void_p), self.n)
        return a

    def close(self):
        if self.h:
            self.lib.lib.sphx_body_destroy(self.h)
            self.h = C.c_void_p()


class CellLinkedList:
    """CellLinkedList<SPHAdaptation> over `bounds` (the system bounds) with grid
    spacing = cut-off radius 2h and a 2-cell buffer by default."""

    def __init__(self, body, bounds_lower, bounds_upper, grid_spacing=None, buffer_width=2):
        self.body, self.lib = body, body.lib
        gs = grid_spacing if grid_spacing is not None else 2.0 * body.smoothing_length
        lo = np.ascontiguousarray(bounds_lower, dtype=np.float64)
        up = np.ascontiguousarray(bounds_upper, dtype=np.float64)
        self.h = C.c_void_p()
        self.lib.call("sphx_cell_linked_list_create", body.h, as_dp(lo), as_dp(up), float(gs),
                      int(buffer_width), C.byref(self.h))
        cells = (C.c_int * 3)()
        lower = (C.c_double * 3)()
        g = C.c_double()
        self.lib.call("sphx_cell_linked_list_mesh", self.h, cells, lower, C.byref(g))
        self.all_cells = tuple(cells[k] for k in range(body.dim))
        self.lower = tuple(lower[k] for k in range(body.dim))
        self.grid_spacing = g.value
        self.n_cells = int(np.prod(self.all_cells))

    def read(self):
        off = np.empty(self.n_cells + 1, dtype=np.uint32)
        idx = np.empty(self.body.n, dtype=np.uint32)
        self.lib.call("sphx_cell_linked_list_read", self.h, as_up(off), off.size, as_up(idx), idx.size)
        return off, idx

    def close(self):
        if self.h:
            self.lib.lib.sphx_cell_linked_list_destroy(self.h)
            self.h = C.c_void_p()


class _Relation:
    def read(self):
        n = self.source.n
        size = np.empty(n, dtype=np.uint32)
        off = np.empty(n + 1, dtype=np.uint32)
        total = C.c_size_t()
        self.lib.call("sphx_relation_read", self.h, as_up(size), as_up(off), None, 0, C.byref(total))
        idx = np.empty(max(1, total.value), dtype=np.uint32)
        self.lib.call("sphx_relation_read", self.h, None, None, as_up(idx), idx.size, C.byref(total))
        return size, off, idx[:total.value]

    def close(self):
        if self.h:
            self.lib.lib.sphx_relation_destroy(self.h)
            self.h = C.c_void_p()


class Inner(_Relation):
    """Inner<> relation of a body (relation_ck.h:89-130)."""

    def __init__(self, body, cll):
        self.source = self.target = body
        self.cll, self.lib = cll, body.lib
        self.h = C.c_void_p()
        self.lib.call("sphx_relation_create_inner", body.h, cll.h, C.byref(self.h))


class Contact(_Relation):
    """Contact<> relation from `source` to one contact body (relation_ck.h:132-174)."""

    def __init__(self, source, target, target_cll):
        self.source, self.target, self.cll, self.lib = source, target, target_cll, source.lib
        self.h = C.c_void_p()
        self.lib.call("sphx_relation_create_contact", source.h, target.h, target_cll.h, C.byref(self.h))


# --------------------------------------------------------------------------
# dynamics
# --------------------------------------------------------------------------
class UpdateCellLinkedList:
    def __init__(self, cll):
        self.cll = cll

    def exec(self, dt=0.0):
        self.cll.lib.call("sphx_update_cell_linked_list", self.cll.h)


class UpdateRelation:
    def __init__(self, *relations):
        self.relations = relations

    def exec(self, dt=0.0):
        for r in self.relations:
            r.lib.call("sphx_update_relation", r.h)


class ParticleSortCK:
    def __init__(self, body, cll):
        self.body, self.cll = body, cll

    def exec(self, dt=0.0):
        self.body.lib.call("sphx_particle_sort", self.body.h, self.cll.h)


class GravityForceCK:
    def __init__(self, body, gravity):
        self.body = body
        self.g = np.ascontiguousarray(gravity, dtype=np.float64)

    def exec(self, dt=0.0):
        self.body.lib.call("sphx_gravity_force", self.body.h, as_dp(self.g))


class DensityRegularization:
    def __init__(self, body, flow=capi.FLOW_FREE_SURFACE):
        self.body, self.flow = body, flow

    def exec(self, dt=0.0):
        self.body.lib.call("sphx_density_regularization", self.body.h, self.flow)


class AdvectionStepSetup:
    def __init__(self, body):
        self.body = body

    def exec(self, dt=0.0):
        self.body.lib.call("sphx_advection_step_setup", self.body.h)


class UpdateParticlePosition:
    def __init__(self, body):
        self.body = body

    def exec(self, dt=0.0):
        self.body.lib.call("sphx_update_particle_position", self.body.h)


class _FluidInteraction:
    def __init__(self, inner, *contacts):
        self.inner, self.contacts = inner, contacts
        self.lib = inner.lib
        self._ptrs, self._keep = capi.SphxLibrary.ptr_array([c.h for c in contacts])


class CompressionSummation(_FluidInteraction):
    """CompressionSummation<Inner<>, Contact<>>."""

    def exec(self, dt=0.0):
        self.lib.call("sphx_compression_summation", self.inner.h, self._ptrs, len(self.contacts))


class AcousticStep1stHalf(_FluidInteraction):
    """AcousticStep1stHalf<Inner<OneLevel, R, K>, Contact<Wall, R, K>> with the kernel
    correction K = NoKernelCorrectionCK (correction=CORRECTION_NONE) or LinearCorrectionCK."""

    _entry = ("sphx_acoustic_step_1st_half", "sphx_acoustic_step_1st_half_corrected")

    def __init__(self, inner, *walls, riemann=capi.RIEMANN_ACOUSTIC, correction=capi.CORRECTION_NONE):
        super().__init__(inner, *walls)
        self.riemann, self.correction = riemann, correction

    def exec(self, dt=0.0):
        self.lib.call(self._entry[1 if self.correction == capi.CORRECTION_LINEAR else 0], self.inner.h,
                      self._ptrs, len(self.contacts), self.riemann, float(dt))


class AcousticStep2ndHalf(AcousticStep1stHalf):
    _entry = ("sphx_acoustic_step_2nd_half", "sphx_acoustic_step_2nd_half_corrected")


class AcousticStep2ndHalfTraversal(AcousticStep1stHalf):
    """AcousticStep2ndHalf with the inner neighbours found by a cell traversal
    instead of read from the stored lists (3D FP32-mixed, uncorrected; the same
    results bit for bit): the alternative design, kept for its comparison."""

    _entry = ("sphx_acoustic_step_2nd_half_traversal", None)

    def __init__(self, inner, *walls, riemann=capi.RIEMANN_ACOUSTIC):
        super().__init__(inner, *walls, riemann=riemann)


def AcousticStep1stHalfWithWallRiemannCK(inner, *walls):
    return AcousticStep1stHalf(inner, *walls, riemann=capi.RIEMANN_ACOUSTIC)


def AcousticStep2ndHalfWithWallRiemannCK(inner, *walls):
    return AcousticStep2ndHalf(inner, *walls, riemann=capi.RIEMANN_ACOUSTIC)


def AcousticStep1stHalfWithWallRiemannCorrectionCK(inner, *walls):
    return AcousticStep1stHalf(inner, *walls, riemann=capi.RIEMANN_ACOUSTIC,
                               correction=capi.CORRECTION_LINEAR)


def AcousticStep2ndHalfWithWallRiemannCorrectionCK(inner, *walls):
    return AcousticStep2ndHalf(inner, *walls, riemann=capi.RIEMANN_ACOUSTIC,
                               correction=capi.CORRECTION_LINEAR)


class LinearCorrectionMatrixComplex(_FluidInteraction):
    """LinearCorrectionMatrix<Inner<WithUpdate>, Contact<>> built with
    DynamicsArgs(inner, alpha) and the contact relations."""

    def __init__(self, inner, *contacts, alpha=0.0):
        super().__init__(inner, *contacts)
        self.alpha = alpha

    def exec(self, dt=0.0):
        self.lib.call("sphx_linear_correction_matrix", self.inner.h, self._ptrs, len(self.contacts),
                      float(self.alpha))


class FreeSurfaceIndicationComplexSpatialTemporalCK(_FluidInteraction):
    """FreeSurfaceIndicationCK<Inner<WithUpdate>, Contact<>>: "Indicator",
    "PositionDivergence" and "PreviousSurfaceIndicator" of the inner body."""

    def exec(self, dt=0.0):
        self.lib.call("sphx_free_surface_indication", self.inner.h, self._ptrs, len(self.contacts))


class ViscousForceCK(_FluidInteraction):
    """ViscousForceCK<Inner<WithUpdate, Viscosity, K>, Contact<Wall, Viscosity, K>> for the
    body's constant viscosity `mu`; adds the change of "ViscousForce" to "ForcePrior"."""

    def __init__(self, inner, *walls, mu, correction=capi.CORRECTION_NONE):
        super().__init__(inner, *walls)
        self.mu, self.correction = mu, correction

    def exec(self, dt=0.0):
        self.lib.call("sphx_viscous_force", self.inner.h, self._ptrs, len(self.contacts), float(self.mu),
                      self.correction)


def ViscousForceWithWallCK(inner, *walls, mu):
    return ViscousForceCK(inner, *walls, mu=mu, correction=capi.CORRECTION_NONE)


class KernelGradientIntegral(_FluidInteraction):
    """KernelGradientIntegral<Inner<K>, Contact<Boundary, K>>: "KernelGradientIntegral"."""

    def __init__(self, inner, *contacts, correction=capi.CORRECTION_NONE):
        super().__init__(inner, *contacts)
        self.correction = correction

    def exec(self, dt=0.0):
        self.lib.call("sphx_kernel_gradient_integral", self.inner.h, self._ptrs, len(self.contacts),
                      self.correction)


def KernelGradientIntegralComplex(inner, *contacts):
    return KernelGradientIntegral(inner, *contacts, correction=capi.CORRECTION_NONE)


def KernelGradientIntegralCorrectedComplex(inner, *contacts):
    return KernelGradientIntegral(inner, *contacts, correction=capi.CORRECTION_LINEAR)


class TransportVelocityCorrectionCK:
    """TransportVelocityCorrectionCK<SPHBody, Limiter[, BulkParticles]>(body, coefficient)."""

    def __init__(self, body, coefficient=0.2, limiter=capi.LIMITER_NONE, bulk_only=False):
        self.body, self.coefficient, self.limiter, self.bulk_only = body, coefficient, limiter, bulk_only

    def exec(self, dt=0.0):
        self.body.lib.call("sphx_transport_velocity_correction", self.body.h, float(self.coefficient),
                           self.limiter, 1 if self.bulk_only else 0)


class AcousticTimeStepCK:
    def __init__(self, body, acoustic_cfl=0.6):
        self.body, self.cfl = body, acoustic_cfl

    def exec(self, dt=0.0):
        v = C.c_double()
        self.body.lib.call("sphx_acoustic_time_step", self.body.h, float(self.cfl), C.byref(v))
        return v.value


class AdvectionTimeStepCK:
    def __init__(self, body, u_ref, advection_cfl=0.25):
        self.body, self.u_ref, self.cfl = body, u_ref, advection_cfl

    def exec(self, dt=0.0):
        v = C.c_double()
        self.body.lib.call("sphx_advection_time_step", self.body.h, float(self.u_ref),
                           float(self.cfl), C.byref(v))
        return v.value


class TotalMechanicalEnergyCK:
    def __init__(self, body, gravity):
        self.body = body
        self.g = np.ascontiguousarray(gravity, dtype=np.float64)

    def exec(self, dt=0.0):
        v = C.c_double()
        self.body.lib.call("sphx_total_mechanical_energy", self.body.h, as_dp(self.g), C.byref(v))
        return v.value


class UpperFrontInAxisDirectionCK:
    def __init__(self, body, axis=0):
        self.body, self.axis = body, axis

    def exec(self, dt=0.0):
        v = C.c_double()
        self.body.lib.call("sphx_upper_front_in_axis", self.body.h, int(self.axis), C.byref(v))
        return v.value


class InterpolationContact:
    """Interpolation<Contact<Real>> of a scalar field, or Interpolation<Contact<Vecd>> of
    "Velocity" (one row of dim values per observer), of the contact body at observers."""

    def __init__(self, observer_contact, field="Pressure"):
        self.rel = observer_contact
        self.fid, self.kind = capi.FIELDS[field]

    def exec(self, dt=0.0):
        n = self.rel.source.n
        if self.kind == "vec":
            out = np.empty((n, self.rel.source.dim), dtype=np.float64)
        else:
            out = np.empty(n, dtype=np.float64)
        self.rel.lib.call("sphx_interpolate_contact", self.rel.h, self.fid, as_dp(out), n)
        return out


class AdvectionStep:
    """The dual-time advection step of the CK dam-break loop behind one call."""

    def __init__(self, inner, walls, cll, u_ref, riemann=capi.RIEMANN_ACOUSTIC,
                 flow=capi.FLOW_FREE_SURFACE, advection_cfl=0.25, acoustic_cfl=0.6,
                 correction=capi.CORRECTION_NONE, correction_alpha=0.5, surface_indication=False):
        self.inner, self.walls, self.cll, self.lib = inner, walls, cll, inner.lib
        self.cfg = capi.AdvectionConfig(u_ref, advection_cfl, acoustic_cfl, riemann, flow, 0,
                                        correction, correction_alpha, 1 if surface_indication else 0)
        self._ptrs, self._keep = capi.SphxLibrary.ptr_array([w.h for w in walls])

    def exec(self, physical_time, sort_now=False):
        self.cfg.sort_now = 1 if sort_now else 0
        t = C.c_double(physical_time)
        n = C.c_uint32()
        self.lib.call("sphx_advection_step", self.inner.h, self._ptrs, len(self.walls), self.cll.h,
                      C.byref(self.cfg), C.byref(t), C.byref(n))
        return t.value, n.value


class BatchAdvection:
    """Many scenes of one device advanced together (sphx_batch_*): one launch
    per stage over all scenes, captured in a HIP graph and replayed once per
    advection step; every scene's result is the one of its own AdvectionStep.
    `scenes` are AdvectionStep objects (their inner relation, walls, cell list
    and configuration)."""

    def __init__(self, lib, scenes, device=0, options=None):
        self.lib, self.scenes = lib, []
        self.h = C.c_void_p()
        lib.call("sphx_batch_create", device, C.byref(self.h))
        if options is not None:
            self.set_options(**options)
        self.time = np.zeros(0)
        self.steps = np.zeros(0, dtype=np.uint32)
        self.status = np.zeros(0, dtype=np.int32)
        for sc in scenes:
            self.add(sc)

    def add(self, scene, time=0.0):
        """Admit one more scene (an AdvectionStep) at physical time `time`."""
        self.lib.call("sphx_batch_add_scene", self.h, scene.inner.h, scene._ptrs, len(scene.walls),
                      scene.cll.h, C.byref(scene.cfg))
        self.scenes.append(scene)
        self.time = np.append(self.time, time)
        self.steps = np.append(self.steps, np.uint32(0))
        self.status = np.append(self.status, np.int32(0))

    def remove(self, index):
        """Release scene `index` from the batch (later indices drop by one)."""
        self.lib.call("sphx_batch_remove_scene", self.h, int(index))
        del self.scenes[index]
        self.time = np.delete(self.time, index)
        self.steps = np.delete(self.steps, index)
        self.status = np.delete(self.status, index)

    def set_options(self, use_graph=1, autotune=1, tile_waves=1, stage_min=0, substep_margin=1):
        o = capi.BatchOptions(use_graph, autotune, tile_waves, stage_min, substep_margin)
        self.lib.call("sphx_batch_set_options", self.h, C.byref(o))

    def exec(self, sort_now=None):
        """One advection step of every scene; returns the per-scene statuses
        (0, or SPHX_E_NONFINITE for a diverged scene)."""
        if not self.scenes:
            return self.status.copy()
        sort = None
        if sort_now is not None:
            flags = np.ascontiguousarray(sort_now, dtype=np.int32)
            sort = flags.ctypes.data_as(C.POINTER(C.c_int))
        st = self.lib.lib.sphx_batch_advection_step(
            self.h, self.time.ctypes.data_as(C.POINTER(C.c_double)),
            self.steps.ctypes.data_as(C.POINTER(C.c_uint32)), sort,
            self.status.ctypes.data_as(C.POINTER(C.c_int)))
        if st not in (0, 4):  # SPHX_E_NONFINITE is reported per scene
            self.lib.check(st, "sphx_batch_advection_step")
        return self.status.copy()

    def tuning(self):
        import json
        n = C.c_size_t()
        self.lib.call("sphx_batch_tuning", self.h, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        self.lib.call("sphx_batch_tuning", self.h, buf, n.value + 1, C.byref(n))
        return json.loads(buf.value.decode())

    def close(self):
        if self.h:
            self.lib.lib.sphx_batch_destroy(self.h)
            self.h = C.c_void_p()
