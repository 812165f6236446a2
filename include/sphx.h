/*
 * sphx.h -- C ABI of the MI355X-native weakly-compressible SPH step.
 *
 * Every entry point below replaces one computing-kernel (CK) dynamics class of
 * SPHinXsys when the case file selects the HIP execution policy instead of
 * par_host.  The reference class each call stands in for is cited as
 * "reference:" with file:line relative to the SPHinXsys source tree.
 *
 * Conventions
 *  - All functions return sphx_status (0 = SPHX_OK).  On failure the message
 *    is available from sphx_last_error() (thread-local).  Nothing calls exit().
 *  - Host arrays are borrowed only for the duration of a call.
 *  - Vectors cross the boundary as AoS doubles (x,y[,z] per particle), the
 *    layout of Eigen's Vecd in a Real=double build; scalars as double; index
 *    fields as uint32.  The library converts to its own device layout.
 *  - Calls on objects of one context are enqueued on that context's stream
 *    in call order; calls that return host data synchronise the stream.
 *    Different contexts may be driven from different host threads.
 */
#ifndef SPHX_H
#define SPHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int sphx_status;
enum {
    SPHX_OK = 0,
    SPHX_E_ARG = 1,       /* invalid argument or inconsistent objects          */
    SPHX_E_HIP = 2,       /* HIP runtime error                                  */
    SPHX_E_CAPACITY = 3,  /* a caller-provided buffer is too small              */
    SPHX_E_NONFINITE = 4, /* a non-finite value was detected in the state       */
    SPHX_E_NOMEM = 5,     /* device allocation failed                           */
    SPHX_E_STATE = 6      /* call out of order (e.g. relation never updated)    */
};

/* Arithmetic of a body.  FP64: every field and every pair interaction in
 * double (the reference's Real=double build).  MIXED: pair interactions,
 * per-particle fields and their accumulators in float; position,
 * displacement, physical time and global mass/energy reductions in double. */
enum { SPHX_FP64 = 0, SPHX_MIXED = 1 };

/* Body kinds (reference: FluidBody / SolidBody / ObserverBody,
 * src/shared/bodies/base_body.h). */
enum { SPHX_BODY_FLUID = 0, SPHX_BODY_SOLID = 1, SPHX_BODY_OBSERVER = 2 };

/* Particle fields.  Vector fields carry `dim` doubles per particle. */
enum {
    SPHX_FIELD_POSITION = 0,          /* "Position"              Vecd  */
    SPHX_FIELD_VOLUME = 1,            /* "VolumetricMeasure"     Real  */
    SPHX_FIELD_MASS = 2,              /* "Mass"                  Real  */
    SPHX_FIELD_DENSITY = 3,           /* "Density"               Real  */
    SPHX_FIELD_PRESSURE = 4,          /* "Pressure"              Real  */
    SPHX_FIELD_VELOCITY = 5,          /* "Velocity"              Vecd  */
    SPHX_FIELD_FORCE = 6,             /* "Force"                 Vecd  */
    SPHX_FIELD_FORCE_PRIOR = 7,       /* "ForcePrior"            Vecd  */
    SPHX_FIELD_DISPLACEMENT = 8,      /* "Displacement"          Vecd  */
    SPHX_FIELD_COMPRESSION = 9,       /* "Compression"           Real  */
    SPHX_FIELD_COMPRESSION_RATE = 10, /* "CompressionRate"       Real  */
    SPHX_FIELD_VOLUME_REF = 11,       /* "VolumetricMeasureRef"  Real  */
    SPHX_FIELD_COMPRESSION_SUM = 12,  /* "CompressionSummation"  Real  */
    SPHX_FIELD_NORMAL = 13,           /* "NormalDirection"       Vecd (solid) */
    SPHX_FIELD_WALL_VELOCITY = 14,    /* "Velocity" of a solid   Vecd  */
    SPHX_FIELD_WALL_ACCELERATION = 15,/* "Acceleration" of solid Vecd  */
    SPHX_FIELD_ORIGINAL_ID = 16,      /* "OriginalID"            uint32 */
    SPHX_FIELD_SORTED_ID = 17,        /* "SortedID"              uint32 */
    SPHX_FIELD_SORT_SEQUENCE = 18,    /* sort keys after the last sort, uint32 */
    SPHX_FIELD_INTERPOLATED = 19,     /* observer result of sphx_interpolate_contact */
    SPHX_FIELD_CORRECTION_MATRIX = 20,/* "LinearCorrectionMatrix" Matd: dim*dim doubles per
                                         particle, row-major; initial value identity */
    SPHX_FIELD_INDICATOR = 21,        /* "Indicator"               int32 (1: free surface) */
    SPHX_FIELD_PREVIOUS_INDICATOR = 22,/* "PreviousSurfaceIndicator" int32, initial 1   */
    SPHX_FIELD_POSITION_DIVERGENCE = 23,/* "PositionDivergence"    Real  */
    SPHX_FIELD_VISCOUS_FORCE = 24,    /* "ViscousForce"          Vecd  */
    SPHX_FIELD_PREVIOUS_VISCOUS_FORCE = 25, /* "PreviousViscousForce" Vecd (evolving) */
    SPHX_FIELD_KERNEL_GRADIENT_INTEGRAL = 26, /* "KernelGradientIntegral" Vecd */
    SPHX_FIELD_COUNT = 27
};

/* Riemann solvers (reference: riemann_solver_ck.h:166-172,
 * acoustic_step_2nd_half.h:180-188). */
enum {
    SPHX_RIEMANN_ACOUSTIC = 0,    /* AcousticRiemannSolverCK (TruncatedLinear, slope 3) */
    SPHX_RIEMANN_NONE = 1,        /* NoRiemannSolverCK                                   */
    SPHX_RIEMANN_DISSIPATIVE = 2  /* DissipativeRiemannSolverCK (NoLimiter)              */
};

/* Density regularization flow types (reference: density_regularization.h:146-184). */
enum { SPHX_FLOW_INTERNAL = 0, SPHX_FLOW_FREE_SURFACE = 1 };

/* Kernel corrections of the acoustic half steps (reference: NoKernelCorrectionCK,
 * LinearCorrectionCK, general_dynamics/kernel_correction_ck.h:153-187). */
enum { SPHX_CORRECTION_NONE = 0, SPHX_CORRECTION_LINEAR = 1 };

/* Limiters of the transport-velocity correction (reference: NoLimiter, TruncatedLinear
 * with its default slope 100, common/common_functors.h:69-94). */
enum { SPHX_LIMITER_NONE = 0, SPHX_LIMITER_TRUNCATED_LINEAR = 1 };

typedef struct sphx_context sphx_context;
typedef struct sphx_body sphx_body;
typedef struct sphx_cell_linked_list sphx_cell_linked_list;
typedef struct sphx_relation sphx_relation;

typedef struct sphx_body_desc {
    int dim;                   /* 2 or 3                                          */
    int precision;             /* SPHX_FP64 or SPHX_MIXED                         */
    int kind;                  /* SPHX_BODY_*                                     */
    uint32_t n_particles;      /* TotalRealParticles                              */
    double spacing_ref;        /* reference particle spacing dp                  */
    double h_spacing_ratio;    /* h / dp (SPHAdaptation default 1.3)             */
    double rho0;               /* reference density (fluid)                       */
    double c0;                 /* reference sound speed (fluid)                   */
} sphx_body_desc;

/* ---- identification ---------------------------------------------------- */
const char *sphx_backend_name(void);     /* "hip-gfx950" for the device library */
const char *sphx_last_error(void);

/* ---- context ------------------------------------------------------------
 * The *_destroy calls wait for their own context's stream only; the device
 * memory and streams they release stay with the library for the next objects
 * of the process (at most 16 GiB of memory per device; beyond that a released
 * block goes back to HIP at once, and all of it does when an allocation of
 * the library runs out of device memory). */
sphx_status sphx_context_create(int device, sphx_context **out);
sphx_status sphx_context_destroy(sphx_context *ctx);
sphx_status sphx_context_synchronize(sphx_context *ctx);

/* ---- bodies and particle state -----------------------------------------
 * reference: BaseParticles variable registry and DiscreteVariable mirrors,
 * src/shared/particles/base_particles.h:109-176,
 * src/shared/common/sphinxsys_variable.h:282-356;
 * initial values as BaseParticles::initializeBasicDiscreteVariables
 * (base_particles.cpp:33-40) and MatterMaterial::initializeMatterVariables
 * (materials/base_material.cpp:35-41): Compression=1, OriginalID=SortedID=i,
 * Density=rho0, VolumetricMeasureRef=Vol at first upload of Vol. */
sphx_status sphx_body_create(sphx_context *ctx, const sphx_body_desc *desc, sphx_body **out);
sphx_status sphx_body_destroy(sphx_body *body);
sphx_status sphx_body_write(sphx_body *body, int field, const void *host, size_t n_particles);
sphx_status sphx_body_read(sphx_body *body, int field, void *host, size_t n_particles);
/* Smoothing length h = h_spacing_ratio * spacing_ref as the body stores it. */
sphx_status sphx_body_smoothing_length(const sphx_body *body, double *h);

/* ---- cell linked list ----------------------------------------------------
 * reference: CellLinkedList<SPHAdaptation> over the system bounds with grid
 * spacing = kernel cut-off radius and a 2-cell buffer,
 * src/shared/meshes/cell_linked_list.cpp:190-197, base_mesh.cpp:6-16,
 * adaptation.cpp:80-85, base_body.cpp:99-105. */
sphx_status sphx_cell_linked_list_create(sphx_body *body, const double *bounds_lower,
                                         const double *bounds_upper, double grid_spacing,
                                         uint32_t buffer_width, sphx_cell_linked_list **out);
sphx_status sphx_cell_linked_list_destroy(sphx_cell_linked_list *cll);
/* all_cells[dim], mesh_lower_bound[dim], grid_spacing. */
sphx_status sphx_cell_linked_list_mesh(const sphx_cell_linked_list *cll, int *all_cells,
                                       double *lower_bound, double *grid_spacing);
/* reference: UpdateCellLinkedList<P, RealBody>::exec,
 * shared_ck/particle_dynamics/configuration_dynamics/update_cell_linked_list.hpp:75-106 */
sphx_status sphx_update_cell_linked_list(sphx_cell_linked_list *cll);
/* No reference counterpart: sphx_update_cell_linked_list for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same result; its workspace (a cursor per cell and the scan's
 * tile sums) is a fixed device array of the library (16 MiB per device, held
 * until the process ends) instead of the context's grow-only scratch, and it
 * records no profiling events.  At most 4,194,304 cells (SPHX_E_CAPACITY
 * beyond).  Every call on one device shares that workspace: calls that
 * overlap on one device (contexts driven from several host threads) race. */
sphx_status sphx_update_cell_linked_list_noalloc(sphx_cell_linked_list *cll);
/* cell_offset[n_cells+1], particle_index[n_particles] (per-cell index order
 * is ascending particle index; the reference fills it with relaxed atomics). */
sphx_status sphx_cell_linked_list_read(sphx_cell_linked_list *cll, uint32_t *cell_offset,
                                       size_t n_cell_offset, uint32_t *particle_index,
                                       size_t n_particle_index);

/* ---- relations -------------------------------------------------------------
 * reference: Inner<> / Contact<> (shared_ck/body_relation/relation_ck.h:89-174)
 * with Neighbor<SPHAdaptation,SPHAdaptation> (neighbor_method.hpp:66-161). */
sphx_status sphx_relation_create_inner(sphx_body *body, sphx_cell_linked_list *cll,
                                       sphx_relation **out);
sphx_status sphx_relation_create_contact(sphx_body *source, sphx_body *target,
                                         sphx_cell_linked_list *target_cll,
                                         sphx_relation **out);
sphx_status sphx_relation_destroy(sphx_relation *rel);
/* reference: UpdateRelation<P, Inner<>>::exec (update_body_relation.hpp:117-164)
 *        and UpdateRelation<P, Contact<>>::exec (update_body_relation.hpp:240-288). */
sphx_status sphx_update_relation(sphx_relation *rel);
/* neighbor_size[n_src], particle_offset[n_src+1], neighbor_index[*total].
 * Pass neighbor_index=NULL to query *total only. */
sphx_status sphx_relation_read(sphx_relation *rel, uint32_t *neighbor_size,
                               uint32_t *particle_offset, uint32_t *neighbor_index,
                               size_t capacity, size_t *total);

/* ---- configuration dynamics -------------------------------------------------
 * reference: ParticleSortCK<P>::exec (configuration_dynamics/particle_sort_ck.hpp:76-104):
 * Morton keys of the CLL mesh cells, sort, permutation of the evolving
 * variables (Position, VolumetricMeasure, OriginalID, Velocity, Mass,
 * ForcePrior, Compression, CompressionRate, VolumetricMeasureRef), SortedID. */
sphx_status sphx_particle_sort(sphx_body *body, sphx_cell_linked_list *cll);

/* ---- state dynamics ------------------------------------------------------- */
/* reference: StateDynamics<P, GravityForceCK<Gravity>> (general_dynamics/force_prior_ck.hpp:37-43) */
sphx_status sphx_gravity_force(sphx_body *fluid, const double *gravity);
/* reference: DensityRegularization<SPHBody, WeaklyCompressibleFluid, FlowType>
 * (fluid_dynamics/density_regularization.hpp:85-118) */
sphx_status sphx_density_regularization(sphx_body *fluid, int flow_type);
/* reference: AdvectionStepSetup (fluid_dynamics/fluid_time_step_ck.h:133-160) */
sphx_status sphx_advection_step_setup(sphx_body *fluid);
/* reference: UpdateParticlePosition (fluid_dynamics/fluid_time_step_ck.h:162-184) */
sphx_status sphx_update_particle_position(sphx_body *fluid);

/* ---- interaction dynamics ------------------------------------------------- */
/* reference: InteractionDynamicsCK<P, CompressionSummation<Inner<>, Contact<>>>
 * (fluid_dynamics/density_regularization.hpp:14-83). contact may be NULL when n_contact=0. */
sphx_status sphx_compression_summation(sphx_relation *inner, sphx_relation *const *contact,
                                       int n_contact);
/* reference: InteractionDynamicsCK<P, AcousticStep1stHalf<Inner<OneLevel,R,NoKernelCorrectionCK>,
 *                                  Contact<Wall,R,NoKernelCorrectionCK>>>
 * (fluid_dynamics/acoustic_step_1st_half.hpp:66-180). walls may be NULL when n_wall=0. */
sphx_status sphx_acoustic_step_1st_half(sphx_relation *inner, sphx_relation *const *walls,
                                        int n_wall, int riemann, double dt);
/* reference: ... AcousticStep2ndHalf<...> (fluid_dynamics/acoustic_step_2nd_half.hpp:33-137). */
sphx_status sphx_acoustic_step_2nd_half(sphx_relation *inner, sphx_relation *const *walls,
                                        int n_wall, int riemann, double dt);
/* No counterpart in the reference: sphx_acoustic_step_2nd_half with the inner neighbours
 * found by a cell traversal (one wave per 64-particle chunk, the chunk's search-box cells
 * staged in LDS, the neighbour criterion evaluated per candidate) instead of read from the
 * stored lists, for the comparison of the two designs. Same results bit for bit (the
 * neighbours are taken in list order); 3D FP32-mixed bodies only (SPHX_E_ARG otherwise).
 * The inner relation's cell linked list must be current. */
sphx_status sphx_acoustic_step_2nd_half_traversal(sphx_relation *inner, sphx_relation *const *walls,
                                                  int n_wall, int riemann, double dt);
/* reference: InteractionDynamicsCK<P, LinearCorrectionMatrixComplex> built with
 * DynamicsArgs(inner, alpha) and the wall contact (general_dynamics/kernel_correction_ck.hpp:37-90,
 * test_3d_dambreak_sycl/dambreak.cpp:123-124): B_i = -sum_j r_ij (x) grad W_ij V_j over the
 * inner and contact neighbours, then B_i = w inverseTikhonov(B_i, 1e-8) + (1 - w) I with
 * w = det / (det + max(alpha - det, 0)).  Result in SPHX_FIELD_CORRECTION_MATRIX. */
sphx_status sphx_linear_correction_matrix(sphx_relation *inner, sphx_relation *const *contact,
                                          int n_contact, double alpha);
/* reference: AcousticStep1stHalfWithWallRiemannCorrectionCK and its Riemann variants,
 * AcousticStep1stHalf<Inner<OneLevel,R,LinearCorrectionCK>, Contact<Wall,R,LinearCorrectionCK>>
 * (acoustic_step_1st_half.h:199-201, .hpp:89-111, 157-180): the pressure force with
 * AverageP(B_j p_i, B_i p_j) inside the fluid and (p_i + p_w) B_i at the wall, B from the
 * last sphx_linear_correction_matrix (or a SPHX_FIELD_CORRECTION_MATRIX write). */
sphx_status sphx_acoustic_step_1st_half_corrected(sphx_relation *inner, sphx_relation *const *walls,
                                                  int n_wall, int riemann, double dt);
/* reference: AcousticStep2ndHalfWithWallRiemannCorrectionCK (acoustic_step_2nd_half.h:189-191,
 * .hpp:53-73, 117-137): velocity divergence with B_i e_ij. */
sphx_status sphx_acoustic_step_2nd_half_corrected(sphx_relation *inner, sphx_relation *const *walls,
                                                  int n_wall, int riemann, double dt);
/* reference: InteractionDynamicsCK<P, FreeSurfaceIndicationComplexSpatialTemporalCK>
 * (general_dynamics/surface_indication/surface_indication_ck.hpp:48-178): position divergence
 * -sum dW_ij V_j r_ij (inner, held above threshold 0.75 dim for a particle that was not and is
 * not next to the previous free surface, then the contact part), Indicator = 1 unless the
 * divergence exceeds the threshold with no surface-side neighbour closer than h;
 * PreviousSurfaceIndicator = Indicator. */
sphx_status sphx_free_surface_indication(sphx_relation *inner, sphx_relation *const *contact,
                                         int n_contact);

/* reference: InteractionDynamicsCK<P, ViscousForceCK<Inner<WithUpdate, Viscosity, K>,
 * Contact<Wall, Viscosity, K>>> (fluid_dynamics/viscous_force.hpp:38-104) with the update
 * of ForcePriorCK (general_dynamics/force_prior_ck.h:50-59): "ViscousForce" from the inner
 * and wall neighbours for the constant viscosity mu (Viscosity material), then
 * ForcePrior += ViscousForce - PreviousViscousForce; PreviousViscousForce = ViscousForce.
 * correction: SPHX_CORRECTION_* (K; ViscousForceWithWallCK uses NONE). */
sphx_status sphx_viscous_force(sphx_relation *inner, sphx_relation *const *walls, int n_wall, double mu,
                               int correction);
/* reference: InteractionDynamicsCK<P, KernelGradientIntegral<Inner<K>, Contact<Boundary, K>>>
 * (general_dynamics/kernel_gradient_integral.hpp:37-49,72-84; KernelGradientIntegralComplex /
 * ...CorrectedComplex): "KernelGradientIntegral" = -sum (B_i + B_j) dW V_j e_ij over the inner
 * neighbours - sum 2 B_i dW V_j e_ij over the contact neighbours. */
sphx_status sphx_kernel_gradient_integral(sphx_relation *inner, sphx_relation *const *contact,
                                          int n_contact, int correction);
/* reference: StateDynamics<P, TransportVelocityCorrectionCK<SPHBody, Limiter[, BulkParticles]>>
 * (fluid_dynamics/transport_velocity_correction_ck.hpp:11-51): Displacement += coefficient h^2
 * limiter(h^2 |g|^2) g for g = KernelGradientIntegral (uniform resolution: h ratio 1); with
 * bulk_only, only particles whose Indicator is 0. coefficient 0.2 by default in the reference. */
sphx_status sphx_transport_velocity_correction(sphx_body *fluid, double coefficient, int limiter,
                                               int bulk_only);

/* ---- reduce dynamics -------------------------------------------------------- */
/* reference: ReduceDynamicsCK<P, AcousticTimeStepCK<WeaklyCompressibleFluid>>
 * (fluid_dynamics/fluid_time_step_ck.hpp:30-57); acoustic_cfl default 0.6. */
sphx_status sphx_acoustic_time_step(sphx_body *fluid, double acoustic_cfl, double *dt);
/* reference: ReduceDynamicsCK<P, AdvectionTimeStepCK> (fluid_time_step_ck.cpp:12-26,
 * fluid_time_step_ck.h:84-120); advection_cfl default 0.25. */
sphx_status sphx_advection_time_step(sphx_body *fluid, double u_ref, double advection_cfl,
                                     double *dt);
/* reference: ReducedQuantityRecording<P, TotalMechanicalEnergyCK>
 * (general_dynamics/general_reduce_ck.h:69-96, external_force.h:53-56). */
sphx_status sphx_total_mechanical_energy(sphx_body *fluid, const double *gravity, double *energy);
/* reference: UpperFrontInAxisDirectionCK (general_dynamics/general_reduce_ck.h:197). */
sphx_status sphx_upper_front_in_axis(sphx_body *fluid, int axis, double *front);

/* ---- observation -------------------------------------------------------------- */
/* reference: Interpolation<Contact<Real>> over an observer contact relation
 * (general_dynamics/interpolation_dynamics.hpp:46-60); field is SPHX_FIELD_PRESSURE,
 * _DENSITY or _COMPRESSION of the contact body (result out[n_observer]) or
 * SPHX_FIELD_VELOCITY (Interpolation<Contact<Vecd>>, result out[n_observer * dim]). */
sphx_status sphx_interpolate_contact(sphx_relation *observer_contact, int field, double *out,
                                     size_t n_observer);

/* ---- dual-time stepping driver ---------------------------------------------------
 * One advection step of the weakly-compressible loop of the CK dam-break
 * cases (tests/tests_sycl/3d_examples/test_3d_dambreak_sycl/dambreak.cpp:188-222):
 * compression summation -> regularization -> setup -> advection dt ->
 * [free-surface indication] -> [linear correction matrix] ->
 * acoustic sub-steps (dt, 1st half, 2nd half) until their sum reaches the
 * advection dt -> position update -> [sort when sort_now] -> cell linked
 * list -> relations.  The advection dt, each sub-step's acoustic dt, the loop
 * condition and the physical time are evaluated on the device in double (the
 * reference's host arithmetic); the host reads the loop state back once per
 * batch of sub-steps (a batch is as long as the previous call's loop).  A
 * non-finite acoustic dt ends the call with SPHX_E_NONFINITE before the
 * position update.  On return *physical_time has advanced and *acoustic_steps holds the number
 * of sub-steps taken. No reference counterpart: it is the case-file loop body
 * executed behind one call. */
typedef struct sphx_advection_config {
    double u_ref;          /* AdvectionTimeStepCK U_ref                 */
    double advection_cfl;  /* 0.25                                        */
    double acoustic_cfl;   /* 0.6                                         */
    int riemann;           /* SPHX_RIEMANN_*                              */
    int flow_type;         /* SPHX_FLOW_*                                 */
    int sort_now;          /* run ParticleSortCK after the position update */
    int correction;        /* SPHX_CORRECTION_*: LINEAR runs LinearCorrectionMatrixComplex
                              before the sub-steps and the corrected half steps         */
    double correction_alpha; /* LinearCorrectionMatrix alpha (0.5 in the CK cases)    */
    int surface_indication; /* run FreeSurfaceIndicationComplexSpatialTemporalCK      */
} sphx_advection_config;

sphx_status sphx_advection_step(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                                sphx_cell_linked_list *fluid_cll,
                                const sphx_advection_config *cfg, double *physical_time,
                                uint32_t *acoustic_steps);

/* ---- batched scenes --------------------------------------------------------------
 * No reference counterpart (the reference advances one scene per process,
 * with one blocking launch per dynamics, implementation_sycl.h:43-92).  A
 * batch advances many independent scenes of one device together: every stage
 * of the advection step is one launch over the blocks of all scenes, which
 * read their own arguments from a scene table on the device (particle and
 * relation pointers, parameters, time-step slots); each scene's dual-time
 * loop runs on the device from its own slots.  The step is captured once in a
 * HIP graph (sphx_batch_options.use_graph) and replayed: one graph launch per
 * advection step for the whole batch.  The graph holds as many acoustic
 * sub-steps as the scenes needed last time plus a margin; a scene that needs
 * more, whose neighbour lists outgrow their width, or that sorts this step is
 * finished by the host through the single-scene path.  An autotuner times
 * launch-shape candidates over the first steps (results do not depend on the
 * shape) and keeps the fastest; sphx_batch_tuning reports what it measured.
 *
 * Every scene is a fluid body with its inner relation, 0..4 wall contact
 * relations and its cell linked list, built and initialised as for
 * sphx_advection_step, on contexts of the batch's device; all scenes share
 * dimension, precision and Riemann solver, and use no kernel correction and
 * no surface indication.  While a batch step runs, nothing else may use the
 * scenes; work issued on a scene's context between batch steps must be
 * synchronised (sphx_context_synchronize) before the next batch step.  Each
 * scene's result is bit-identical to the same scene advanced alone with
 * sphx_advection_step. */
typedef struct sphx_batch sphx_batch;

typedef struct sphx_batch_options {
    int use_graph;          /* 1 (default): capture the step in a HIP graph and replay it;
                               0: the same launches issued directly every step          */
    int autotune;           /* 1 (default): time the shape candidates over the first
                               steps and keep the fastest; 0: use the fields below     */
    int tile_waves;         /* waves per workgroup of the tiled neighbour build: 1, 2, 4 */
    uint32_t stage_min;     /* scenes with at least this many fluid particles run the
                               acoustic passes from LDS-staged neighbour records       */
    uint32_t substep_margin;/* captured sub-steps beyond the last step's largest count */
} sphx_batch_options;

sphx_status sphx_batch_create(int device, sphx_batch **out);
sphx_status sphx_batch_destroy(sphx_batch *batch);
sphx_status sphx_batch_set_options(sphx_batch *batch, const sphx_batch_options *options);
/* adds a scene (its index is the number of scenes added before it) */
sphx_status sphx_batch_add_scene(sphx_batch *batch, sphx_relation *inner, sphx_relation *const *walls,
                                 int n_wall, sphx_cell_linked_list *fluid_cll,
                                 const sphx_advection_config *cfg);
/* removes a scene (the indices of the scenes after it drop by one); the
 * scene's objects are untouched and may be used alone again */
sphx_status sphx_batch_remove_scene(sphx_batch *batch, uint32_t index);
/* One advection step of every scene (as sphx_advection_step; cfg->sort_now is
 * replaced by sort_now[scene], null: no sort).  physical_time[scene] in/out,
 * acoustic_steps[scene] out, status[scene] out (may be null).  A scene whose
 * state diverged reports SPHX_E_NONFINITE, stays in the batch without
 * advancing, and makes the call return SPHX_E_NONFINITE after the others have
 * advanced. */
sphx_status sphx_batch_advection_step(sphx_batch *batch, double *physical_time, uint32_t *acoustic_steps,
                                      const int *sort_now, int *status);
/* JSON text of the tuning (candidates, cost per particle sub-step, choice) and
 * of the launch counts; *length receives the full length (without the NUL). */
sphx_status sphx_batch_tuning(const sphx_batch *batch, char *json, size_t capacity, size_t *length);

/* ---- instrumentation ------------------------------------------------------------
 * Kernel timing with HIP events recorded on the context's stream around each
 * launch of the named stage (no reference counterpart; the reference times
 * whole phases with TickCount, dambreak_sycl.cpp:175-259).  Stage names:
 * "acoustic_pass1", "acoustic_pass2", "acoustic_init", "compression_summation",
 * "relation_build" (tiled list build of inner relations), "relation_build_contact"
 * (tiled list build of contact relations), "relation_scan" (per-lane scan build,
 * used when a list outgrows the LDS tile), "relation_stage" (LDS staging sets
 * of the FP32-mixed inner relation), "cell_linked_list", "particle_sort",
 * "time_step_reduce", "state_update", "linear_correction" (correction matrix and
 * its neighbour sum), "surface_indication".  Reading synchronises the stream and
 * returns the accumulated milliseconds and launch count since enabling.
 * enable: 0 off, 1 every stage, 2 only "acoustic_pass1" and "acoustic_pass2"
 * (each event pair delays the next launch by a few microseconds, so timing
 * every small kernel of a step slows the step by 2-3 %). */
sphx_status sphx_profile_enable(sphx_context *ctx, int enable);
sphx_status sphx_profile_read(sphx_context *ctx, const char *stage, double *total_ms,
                              uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* SPHX_H */
