// sphx_internal.hpp -- host-side objects behind the opaque handles of sphx.h
// and the launch functions each kernel translation unit provides.
#pragma once
#include "../../include/sphx.h"
#include "sphx_device.hpp"

#include <map>
#include <string>
#include <vector>

namespace sphx
{
void set_error(const std::string &msg);
sphx_status hip_fail(hipError_t e, const char *what);

#define SPHX_HIP(call)                                \
    do                                                \
    {                                                 \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess)                         \
            return ::sphx::hip_fail(e_, #call);       \
    } while (0)

// Device memory of the objects of a context (s is the context's stream).
// dev_free waits for that stream only and keeps the block for a later
// dev_alloc of the process (up to the block's size and at least half of it)
// instead of returning it with hipFree, which waits for every stream of the
// device: with many contexts running side by side, each scene's set-up and
// teardown would otherwise stall all the others.  Kept blocks are returned to
// HIP when an allocation runs out of memory.
sphx_status dev_alloc(hipStream_t s, void **p, size_t bytes);
void dev_free(hipStream_t s, void *p);

// grow-only device scratch owned by a context (and used on its stream only)
struct Scratch
{
    void *ptr = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    sphx_status reserve(size_t n);
    void release();
};

constexpr int kScalarSlots = 1024;
constexpr uint32_t kStageBlock = 256;          // source particles per staging block
constexpr uint32_t kStageOverflow = 0xffffffffu;
constexpr uint32_t kStageCap = 1536;           // distinct neighbours staged per block (LDS records)
// reported through the relation build's `needed` slot when a search box holds
// more candidates than the tiled build's masks: rebuild with the scan kernel
constexpr unsigned long long kTooDense = 1ull << 40;
enum ScalarSlot
{
    SLOT_TOTAL = 2,      // u32 totals returned by scans (stored as u64)
    SLOT_NEEDED = 4,     // u64: neighbour count beyond the slice width
    SLOT_SUM_BASE = 8,   // double partial sums region start (reduce_sum final)
    SLOT_STEP = 16,      // device-side sub-step control of sphx_advection_step (StepSlot offsets)
    SLOT_MAX = 512       // kMaxLanes ordered-bits max accumulators, kMaxStride apart
};
// offsets from SLOT_STEP: the dual-time loop state kept on the device so that
// the acoustic sub-steps of an advection step run without host round trips
enum StepSlot
{
    STEP_ADV_DT = 0,     // double: advection dt of the step
    STEP_DT = 1,         // double: acoustic dt of the current sub-step
    STEP_RELAXATION = 2, // double: relaxation time so far
    STEP_TIME = 3,       // double: physical time
    STEP_COUNT = 4,      // u64: sub-steps run
    STEP_ACTIVE = 5,     // u64: 1 while the current sub-step runs
    STEP_ERROR = 6,      // u64: 1 when a non-finite acoustic dt stopped the loop
    STEP_SLOTS = 7
};
} // namespace sphx

struct sphx_context
{
    int device = 0;
    hipStream_t stream = nullptr;
    size_t max_lds = 65536; // shared memory per workgroup of the device
    // HIP-event kernel timing (sphx_profile_*)
    bool profiling = false;
    bool profile_passes_only = false; // time only the acoustic pass launches
    struct PendingEvent
    {
        std::string stage;
        hipEvent_t start, stop;
    };
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> event_pool; // recorded events ready for reuse
    std::map<std::string, std::pair<double, uint64_t>> profile; // stage -> (ms, launches)
    unsigned long long *scalars_d = nullptr; // kScalarSlots u64 on the device
    unsigned long long *scalars_h = nullptr; // pinned mirror
    sphx::Scratch sort_keys, sort_vals, sort_hist, scan_blocks, reduce_parts, gather_tmp, cll_cursor;
    // while set, the acoustic kernels take dt and their run flag from the SLOT_STEP slots
    bool device_step = false;
    uint32_t last_substeps = 0; // sub-steps of the previous sphx_advection_step
};

struct sphx_body
{
    sphx_context *ctx = nullptr;
    sphx_body_desc desc{};
    int D = 3, P = 0;
    uint32_t n = 0;
    double h = 0.0;      // smoothing length (computed in double as the FP64 reference)
    bool vol_written = false;
    uint64_t config_version = 0; // bumped whenever positions or particle order change
    size_t tsize = 8;    // sizeof(T)
    double *pos[3] = {nullptr, nullptr, nullptr};
    double *dpos[3] = {nullptr, nullptr, nullptr};
    void *posvol = nullptr; // Rec4<T>
    void *vel = nullptr;    // Rec4<T>
    sphx::Rec4<float> *pos_f32 = nullptr; // float (x, y, z|., .) for the neighbour pre-filter (aliases posvol in MIXED)
    void *force[3] = {nullptr, nullptr, nullptr};
    void *force_prior[3] = {nullptr, nullptr, nullptr};
    void *normal[3] = {nullptr, nullptr, nullptr};
    void *wall_vel[3] = {nullptr, nullptr, nullptr};
    void *wall_acc[3] = {nullptr, nullptr, nullptr};
    void *Vol = nullptr, *mass = nullptr, *rho = nullptr, *p = nullptr, *C = nullptr;
    void *dC = nullptr, *Vol_ref = nullptr, *S = nullptr, *interp = nullptr;
    uint32_t *original_id = nullptr, *sorted_id = nullptr, *sequence = nullptr, *perm = nullptr;
    // "LinearCorrectionMatrix": component (a, b) in corr[a * D + b]; and, for the
    // corrected pressure force, grad_sum = sum_j B_j e_ij dW_ij V_j over the inner
    // neighbours, valid for the state stamped in grad_key
    void *corr[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void *grad_sum[3] = {nullptr, nullptr, nullptr};
    struct GradKey
    {
        uint64_t state = ~0ull;
        const sphx_relation *rel = nullptr;
        uint64_t builds = 0;
    } grad_key;
    // bumped whenever positions, volumes or the correction matrices change
    uint64_t state_version = 0;
    // "Indicator", "PreviousSurfaceIndicator", "PositionDivergence"
    int32_t *indicator = nullptr, *prev_indicator = nullptr;
    void *pos_div = nullptr;
    // "ViscousForce", "PreviousViscousForce" (evolving), "KernelGradientIntegral"
    void *visc[3] = {nullptr, nullptr, nullptr};
    void *prev_visc[3] = {nullptr, nullptr, nullptr};
    void *kgi[3] = {nullptr, nullptr, nullptr};
    std::vector<void *> allocations;
};

struct sphx_cell_linked_list
{
    sphx_body *body = nullptr;
    sphx::MeshParams mesh{};
    uint64_t ncells = 0;
    int key_bits = 0;
    uint32_t *cell_offset = nullptr;    // ncells + 1
    uint32_t *particle_index = nullptr; // n
    uint32_t *keys = nullptr;           // n (sort keys scratch)
    bool built = false;
    uint64_t built_version = 0;         // body config_version the list was built for
    uint64_t build_count = 0;           // updates so far
    uint8_t *near = nullptr;            // per cell: a particle within the +-1 search box (contact targets)
    uint64_t near_build = 0;            // build_count the flags were computed for
};

struct sphx_relation
{
    bool inner = true;
    sphx_body *source = nullptr, *target = nullptr;
    sphx_cell_linked_list *cll = nullptr; // of the target
    uint32_t *neighbor_size = nullptr;    // n_src + 1 (last entry stays 0)
    uint32_t *particle_offset = nullptr;  // n_src + 1, CSR offsets of the reference layout
    uint32_t *neighbor_index = nullptr;   // CSR NeighborIndex, built on read only
    size_t capacity = 0, total = 0;
    // device layout used by the kernels: 64-wide slices, entry k of particle i
    // at slice[chunk_offset[i / 64] + 64 k + i % 64]
    uint32_t *chunk_offset = nullptr;     // n_chunks + 1
    uint32_t *slice = nullptr;
    size_t slice_capacity = 0, slice_total = 0;
    uint32_t kmax = 0;                    // rows per 64-particle chunk of the slices
    uint32_t chunk_kmax = 0;              // slice width chunk_offset was last written for
    bool csr_valid = false;
    // LDS staging of the inner relation: per block of kStageBlock source
    // particles, the set of distinct neighbours (stage_keys[b * stage_cap + r],
    // stage_count[b] of them, or kStageOverflow) and the neighbour entries as
    // 16-bit ranks into that set (local_slice, same slice layout)
    bool staged = false;
    bool stage_stale = false;             // lists rebuilt without their staging sets (the passes gather directly)
    uint32_t stage_cap = 0;
    uint32_t *stage_count = nullptr;
    uint32_t *stage_keys = nullptr;
    uint16_t *local_slice = nullptr;
    size_t local_capacity = 0;
    bool built = false;
    uint64_t builds = 0;                  // completed list builds
    uint64_t built_version = 0, target_version = 0; // config versions the lists were built for
    double inv_h = 0, src_inv_h = 0;      // Neighbor<SPHAdaptation,SPHAdaptation> constants
    void *poly = nullptr;                 // KernelPoly<T> on the device
};

namespace sphx
{
// ---- configuration kernels (config_kernels.hip) ----
// Caller-provided device workspace of the cell-list update (null: the
// context's grow-only scratch is used and grown with dev_alloc): a cursor
// per cell and the tile sums of the cell-offset scan (words of 4 B).
struct CllWork
{
    uint32_t *cursor = nullptr;
    uint32_t *scan_sums = nullptr;
};
sphx_status exclusive_scan_u32(sphx_context *ctx, const uint32_t *in, uint32_t *out, size_t n,
                               uint32_t *total_host, uint32_t *sums = nullptr);
sphx_status radix_sort_pairs(sphx_context *ctx, uint32_t *keys, uint32_t *vals, size_t n, int bits);
sphx_status cell_linked_list_update(sphx_cell_linked_list *cll, const CllWork *work = nullptr);
// The allocation-free cell-list update: the cell limit of its file-scope
// device workspace, its per-device address (looked up once per device by
// noalloc_scratch_init, with that device current) and the update itself
constexpr uint64_t kNoAllocMaxCells = 1ull << 22;
sphx_status noalloc_scratch_init(int device);
sphx_status cell_linked_list_update_noalloc(sphx_cell_linked_list *cll);
sphx_status relation_to_csr(sphx_relation *rel);
// squared-distance bounds of the neighbour build's float pre-filter (below
// thr_lo a neighbour, at or above thr_hi not, the double criterion between)
// and the coordinate bound |x| <= pre_xmax under which they hold
void relation_prefilter(const sphx_relation *rel, float *thr_lo, float *thr_hi, double *pre_xmax);
// scan = true forces the per-lane scan kernel (no LDS tile)
sphx_status relation_build(sphx_relation *rel, uint32_t kmax, unsigned long long *needed, bool scan);
// the LDS staging sets of a staged relation, from its finished slices
sphx_status relation_stage(sphx_relation *rel);
// per-cell flags of a target cell list: any particle within the search box
sphx_status cell_near_flags(sphx_cell_linked_list *cll);
sphx_status morton_sequence(sphx_body *b, sphx_cell_linked_list *cll);
sphx_status gather_u32(sphx_context *ctx, uint32_t *field, const uint32_t *perm, size_t n);
sphx_status gather_real(sphx_body *b, void **field, const uint32_t *perm);
sphx_status gather_double(sphx_body *b, double **field, const uint32_t *perm);
sphx_status gather_rec4(sphx_body *b, void **field, const uint32_t *perm);
sphx_status sorted_id_update(sphx_body *b);
sphx_status fill_iota(sphx_context *ctx, uint32_t *a, size_t n);

// ---- fluid kernels (fluid_kernels.hip) ----
sphx_status refresh_posvol(sphx_body *b);
sphx_status gravity_force(sphx_body *b, const double *g);
sphx_status density_regularization(sphx_body *b, int flow);
sphx_status advection_step_setup(sphx_body *b);
sphx_status update_position(sphx_body *b);
sphx_status compression_summation(sphx_relation *inner, sphx_relation *const *c, int nc);
// fuse_init_2nd: pass 1 also runs AcousticStep2ndHalf's initialize
sphx_status acoustic_1st_half(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                              double dt, bool fuse_init_2nd);
// run_init: launch the 2nd half's initialize (false when pass 1 fused it);
// device-side dual-time loop: the step state from the advection speed max in
// the SLOT_MAX lanes, then per sub-step its dt from the signal max there
sphx_status step_begin(sphx_context *ctx, double u_ref, double advection_cfl, double h, double time);
sphx_status substep_begin(sphx_context *ctx, double acoustic_cfl, double h);
// fuse_cfl: max-reduce the next sub-step's signal speed into the SLOT_MAX lanes
sphx_status acoustic_2nd_half(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                              double dt, bool run_init, bool fuse_cfl);
// the same with the inner neighbours found by a cell traversal (3D FP32-mixed)
sphx_status acoustic_2nd_half_traversal(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                                        double dt);
// LinearCorrectionCK variants of the two passes (the 1st refreshes grad_sum when stale)
sphx_status acoustic_1st_half_corrected(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                                        double dt, bool fuse_init_2nd);
sphx_status acoustic_2nd_half_corrected(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                                        double dt, bool run_init, bool fuse_cfl);
sphx_status linear_correction_matrix(sphx_relation *inner, sphx_relation *const *c, int nc, double alpha);
sphx_status free_surface_indication(sphx_relation *inner, sphx_relation *const *c, int nc);
sphx_status viscous_force(sphx_relation *inner, sphx_relation *const *w, int nw, double mu, bool corrected);
sphx_status kernel_gradient_integral(sphx_relation *inner, sphx_relation *const *c, int nc, bool corrected);
sphx_status transport_velocity_correction(sphx_body *b, double coefficient, int limiter, bool bulk_only);
sphx_status acoustic_signal_max(sphx_body *b, unsigned long long *slot);
sphx_status advection_speed_max(sphx_body *b, unsigned long long *slot);
sphx_status acoustic_dt_from_max(sphx_body *b, double cfl, const unsigned long long *slot, double *dt_out);
sphx_status axis_max(sphx_body *b, int axis, unsigned long long *slot);
sphx_status mechanical_energy(sphx_body *b, const double *g, double *out_device);
// component: the velocity component for SPHX_FIELD_VELOCITY
sphx_status interpolate_contact(sphx_relation *rel, int field, int component);
sphx_status fill_u64(sphx_context *ctx, unsigned long long *p, size_t n, unsigned long long v);
sphx_status build_kernel_poly(sphx_relation *rel);

// ---- batched launches over the scenes of a batch (fluid_kernels.hip) ----
// A batch's scene table holds one record per scene (batch_scene_args_bytes
// each, written by batch_scene_args); map[b] = scene << 16 | block of that
// scene over the 256-particle blocks of every scene.
struct BatchFluid
{
    int D = 2, P = 0, riemann = 0;
    hipStream_t stream = nullptr;
    const void *tab = nullptr;
    uint32_t n_scenes = 0;
    const uint32_t *map = nullptr;
    uint32_t blocks = 0;
    uint32_t stage_cap = 0;            // staging capacity of the staged scenes (0: none staged)
    const uint32_t *tail_on = nullptr; // per scene: the device runs the step's tail
    uint32_t *gate = nullptr;          // per scene: sub-steps done and tail on (written by the gate)
    unsigned long long *state = nullptr; // per scene: STEP_SLOTS copies of the step state
};
size_t batch_scene_args_bytes(int P);
// time_dev: the scene's physical time at the start of the step (device);
// staged false: the passes gather from global memory (the staging sets unused)
void batch_scene_args(void *dst, sphx_relation *inner, sphx_relation *const *w, int nw,
                      const sphx_advection_config *cfg, const double *time_dev, bool staged);
// compression summation, regularization + setup, advection dt, first acoustic signal max
sphx_status batch_step_head(const BatchFluid &b);
// one acoustic sub-step: its dt and loop flag, 1st half (with the 2nd half's initialize), 2nd half + CFL
sphx_status batch_substep(const BatchFluid &b);
sphx_status batch_step_gate(const BatchFluid &b);
sphx_status batch_position(const BatchFluid &b);

// ---- batched configuration update (config_kernels.hip) ----
// Per scene: cell list of the fluid (keys, counts, scan, scatter, in-cell
// order); per relation: the one-pass tiled build and, for staged inner
// relations, the staging sets.  Gated per scene like the fluid tail.
struct BatchConfig
{
    int D = 2;
    hipStream_t stream = nullptr;
    const uint32_t *gate = nullptr;
    // cell lists: table, 256-particle map (as BatchFluid), 256-cell map over
    // ncells + 1, scan-tile map over ncells + 1
    const void *cll = nullptr;
    uint32_t n_scenes = 0;
    const uint32_t *map_particles = nullptr;
    uint32_t blocks_particles = 0;
    const uint32_t *map_cells = nullptr;
    uint32_t blocks_cells = 0;
    const uint32_t *map_tiles = nullptr;
    uint32_t blocks_tiles = 0;
    // relations: build table, 64-particle chunk map (relation << 16 | chunk), scene of each relation
    const void *build = nullptr;
    uint32_t n_rel = 0;
    const uint32_t *map_chunks = nullptr;
    uint32_t blocks_chunks = 0;
    const uint32_t *rel_scene = nullptr;
    unsigned long long *needed = nullptr; // per relation, zeroed here before the build
    int tile_waves = 1;                   // waves per workgroup of the tiled build (1, 2 or 4)
    // staging: table per scene (n_src 0: not staged), 256-particle map
    const void *stage = nullptr;
};
size_t batch_cll_args_bytes();
// cursor: ncells words, tile_sums: scan tiles of ncells + 1 (device, owned by the batch)
void batch_cll_args(void *dst, sphx_cell_linked_list *cll, uint32_t *cursor, uint32_t *tile_sums);
uint32_t batch_cll_tiles(const sphx_cell_linked_list *cll);
size_t batch_build_args_bytes();
void batch_build_args(void *dst, sphx_relation *rel, unsigned long long *needed);
size_t batch_stage_args_bytes();
void batch_stage_args(void *dst, sphx_relation *rel); // rel null or not staged: an empty record
sphx_status batch_cell_linked_list(const BatchConfig &b);
sphx_status batch_relations(const BatchConfig &b);

// ---- pieces of the advection step driver (sphx_api.hip) ----
bool relation_ready(sphx_relation *r);
sphx_status check_fluid_relations(sphx_relation *inner, sphx_relation *const *walls, int n_wall, bool need_built);
// the device-side sub-step loop from the current step state until it ends;
// the state is then in the context's pinned mirror
sphx_status advection_substeps(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                               const sphx_advection_config *cfg, uint32_t batch);
// position update, [ParticleSortCK], cell linked list, relations
sphx_status advection_tail(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                           sphx_cell_linked_list *cll, bool sort_now);

// ---- scalar read-back helper (sphx_api.hip) ----
sphx_status read_scalars(sphx_context *ctx, int first, int count);

// records HIP events around the launches issued in its scope when the
// context is profiling
struct ProfileScope
{
    sphx_context *ctx;
    const char *stage;
    hipEvent_t start = nullptr;
    ProfileScope(sphx_context *c, const char *s);
    ~ProfileScope();
};
} // namespace sphx
