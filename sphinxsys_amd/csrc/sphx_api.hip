// sphx_api.hip -- the sphx C ABI (include/sphx.h) on HIP for gfx950.
//
// Object lifetime, argument validation, host<->device layout conversion and
// the order of kernel launches that restates each CK dynamics' exec().
// There is no CPU fallback: every compute entry point launches device code.
#include "sphx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>

namespace sphx
{
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
sphx_status hip_fail(hipError_t e, const char *what)
{
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? SPHX_E_NOMEM : SPHX_E_HIP;
}
static sphx_status arg_error(const std::string &msg)
{
    g_error = msg;
    return SPHX_E_ARG;
}
static sphx_status state_error(const std::string &msg)
{
    g_error = msg;
    return SPHX_E_STATE;
}

// blocks released by destroyed objects, kept for reuse: a block serves a
// request of up to its size and at least half of it
namespace
{
std::mutex g_mem_mutex;
std::map<int, std::multimap<size_t, void *>> g_free_blocks; // device -> (bytes, block)
std::map<void *, std::pair<int, size_t>> g_block_size;      // block -> (device, bytes)
std::map<int, size_t> g_kept_bytes;                         // device -> bytes in g_free_blocks
// at most this much released memory is kept per device; beyond it a released
// block goes back to HIP at once (hipFree, which waits for the device)
constexpr size_t kKeepCap = 16ull << 30;
} // namespace

sphx_status dev_alloc(hipStream_t s, void **p, size_t bytes)
{
    bytes = std::max<size_t>(bytes, 256);
    int dev = 0;
    SPHX_HIP(hipStreamGetDevice(s, &dev));
    {
        std::lock_guard<std::mutex> lock(g_mem_mutex);
        auto &blocks = g_free_blocks[dev];
        auto it = blocks.lower_bound(bytes);
        if (it != blocks.end() && it->first / 2 <= bytes)
        {
            *p = it->second;
            g_kept_bytes[dev] -= it->first;
            blocks.erase(it);
            return SPHX_OK;
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory)
    {
        // return the kept blocks of this device and try once more
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lock(g_mem_mutex);
        for (auto &kv : g_free_blocks[dev])
        {
            (void)hipFree(kv.second);
            g_block_size.erase(kv.second);
        }
        g_free_blocks[dev].clear();
        g_kept_bytes[dev] = 0;
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess)
        return hip_fail(e, "hipMalloc");
    std::lock_guard<std::mutex> lock(g_mem_mutex);
    g_block_size[*p] = {dev, bytes};
    return SPHX_OK;
}
void dev_free(hipStream_t s, void *p)
{
    if (!p)
        return;
    // the block may still be read by kernels queued on its context's stream
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lock(g_mem_mutex);
    auto it = g_block_size.find(p);
    if (it == g_block_size.end())
        return;
    const int dev = it->second.first;
    const size_t bytes = it->second.second;
    if (g_kept_bytes[dev] + bytes > kKeepCap)
    {
        (void)hipFree(p);
        g_block_size.erase(it);
        return;
    }
    g_kept_bytes[dev] += bytes;
    g_free_blocks[dev].emplace(bytes, p);
}

// streams and pinned scalar mirrors of destroyed contexts, reused by the next
// ones (destroying them can wait for the whole device)
namespace
{
std::mutex g_reuse_mutex;
std::map<int, std::vector<hipStream_t>> g_free_streams;
std::vector<unsigned long long *> g_free_mirrors;
} // namespace

sphx_status Scratch::reserve(size_t n)
{
    if (n <= bytes)
        return SPHX_OK;
    // the old buffer is released after the kernels already queued on the stream
    dev_free(stream, ptr);
    ptr = nullptr;
    bytes = 0;
    size_t want = n + n / 4;
    sphx_status s = dev_alloc(stream, &ptr, want);
    if (s)
        return s;
    bytes = want;
    return SPHX_OK;
}
void Scratch::release()
{
    dev_free(stream, ptr);
    ptr = nullptr;
    bytes = 0;
}

sphx_status read_scalars(sphx_context *ctx, int first, int count)
{
    SPHX_HIP(hipMemcpyAsync(ctx->scalars_h + first, ctx->scalars_d + first, count * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, ctx->stream));
    SPHX_HIP(hipStreamSynchronize(ctx->stream));
    return SPHX_OK;
}

template <class T>
sphx_status transfer_scalar_values(sphx_body *b, void *dev, double *h, bool write)
{
    const size_t n = b->n;
    hipStream_t s = b->ctx->stream;
    std::vector<T> tmp(n);
    if (write)
    {
        for (size_t i = 0; i < n; ++i)
            tmp[i] = T(h[i]);
        SPHX_HIP(hipMemcpyAsync(dev, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
        SPHX_HIP(hipStreamSynchronize(s));
    }
    else
    {
        SPHX_HIP(hipMemcpyAsync(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s));
        SPHX_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < n; ++i)
            h[i] = (double)tmp[i];
    }
    return SPHX_OK;
}

// events come from a per-context pool (creating two events per launch costs
// more host time than the small kernels of a step take on the device)
static hipEvent_t pooled_event(sphx_context *ctx)
{
    if (!ctx->event_pool.empty())
    {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

ProfileScope::ProfileScope(sphx_context *c, const char *s) : ctx(c), stage(s)
{
    // a null stage records nothing
    const bool timed = s && ctx->profiling && (!ctx->profile_passes_only || std::strncmp(s, "acoustic_pass", 13) == 0);
    start = timed ? pooled_event(ctx) : nullptr;
    if (start)
        (void)hipEventRecord(start, ctx->stream);
}
ProfileScope::~ProfileScope()
{
    if (!start)
        return;
    hipEvent_t stop = pooled_event(ctx);
    if (!stop)
    {
        ctx->event_pool.push_back(start);
        return;
    }
    (void)hipEventRecord(stop, ctx->stream);
    ctx->pending.push_back({stage, start, stop});
}

static void drain_profile(sphx_context *ctx)
{
    for (auto &e : ctx->pending)
    {
        float ms = 0.0f;
        if (hipEventSynchronize(e.stop) == hipSuccess && hipEventElapsedTime(&ms, e.start, e.stop) == hipSuccess)
        {
            auto &t = ctx->profile[e.stage];
            t.first += ms;
            t.second += 1;
        }
        ctx->event_pool.push_back(e.start);
        ctx->event_pool.push_back(e.stop);
    }
    ctx->pending.clear();
}

namespace
{
template <class T>
__global__ void k_fill(T *a, size_t n, T v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = v;
}

template <class T>
sphx_status fill(sphx_context *ctx, void *p, size_t n, T v)
{
    if (n == 0)
        return SPHX_OK;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (T *)p, n, v);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

sphx_status alloc(sphx_body *b, void **p, size_t bytes)
{
    sphx_status s = dev_alloc(b->ctx->stream, p, bytes);
    if (s)
        return s;
    b->allocations.push_back(*p);
    return SPHX_OK;
}


// host <-> device transfer of one field with layout/precision conversion
template <class T>
sphx_status transfer(sphx_body *b, int field, void *host, bool write)
{
    const size_t n = b->n;
    const int D = b->D;
    hipStream_t s = b->ctx->stream;
    double *h = (double *)host;
    auto scalar_dev = [&](void *dev) -> sphx_status {
        std::vector<T> tmp(n);
        if (write)
        {
            for (size_t i = 0; i < n; ++i)
                tmp[i] = T(h[i]);
            SPHX_HIP(hipMemcpyAsync(dev, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
            SPHX_HIP(hipStreamSynchronize(s));
        }
        else
        {
            SPHX_HIP(hipMemcpyAsync(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s));
            SPHX_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i)
                h[i] = (double)tmp[i];
        }
        return SPHX_OK;
    };
    auto soa_dev = [&](void **dev, auto zero) -> sphx_status {
        using E = decltype(zero);
        std::vector<E> tmp(n);
        for (int k = 0; k < D; ++k)
        {
            if (write)
            {
                for (size_t i = 0; i < n; ++i)
                    tmp[i] = E(h[i * D + k]);
                SPHX_HIP(hipMemcpyAsync(dev[k], tmp.data(), n * sizeof(E), hipMemcpyHostToDevice, s));
                SPHX_HIP(hipStreamSynchronize(s));
            }
            else
            {
                SPHX_HIP(hipMemcpyAsync(tmp.data(), dev[k], n * sizeof(E), hipMemcpyDeviceToHost, s));
                SPHX_HIP(hipStreamSynchronize(s));
                for (size_t i = 0; i < n; ++i)
                    h[i * D + k] = (double)tmp[i];
            }
        }
        return SPHX_OK;
    };
    auto u32_dev = [&](uint32_t *dev) -> sphx_status {
        if (write)
            SPHX_HIP(hipMemcpyAsync(dev, host, n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        else
            SPHX_HIP(hipMemcpyAsync(host, dev, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        SPHX_HIP(hipStreamSynchronize(s));
        return SPHX_OK;
    };
    const bool cm1 = b->P == P_MIXED; // MIXED stores C - 1
    switch (field)
    {
    case SPHX_FIELD_POSITION:
        return soa_dev((void **)b->pos, 0.0);
    case SPHX_FIELD_DISPLACEMENT:
        return soa_dev((void **)b->dpos, 0.0);
    case SPHX_FIELD_VOLUME:
        return scalar_dev(b->Vol);
    case SPHX_FIELD_MASS:
        return scalar_dev(b->mass);
    case SPHX_FIELD_DENSITY:
        return scalar_dev(b->rho);
    case SPHX_FIELD_PRESSURE:
        return scalar_dev(b->p);
    case SPHX_FIELD_COMPRESSION:
    {
        std::vector<double> tmp(n);
        if (write)
        {
            for (size_t i = 0; i < n; ++i)
                tmp[i] = cm1 ? h[i] - 1.0 : h[i];
            return transfer_scalar_values<T>(b, b->C, tmp.data(), true);
        }
        sphx_status st = transfer_scalar_values<T>(b, b->C, tmp.data(), false);
        for (size_t i = 0; i < n; ++i)
            h[i] = cm1 ? tmp[i] + 1.0 : tmp[i];
        return st;
    }
    case SPHX_FIELD_COMPRESSION_RATE:
        return scalar_dev(b->dC);
    case SPHX_FIELD_VOLUME_REF:
        return scalar_dev(b->Vol_ref);
    case SPHX_FIELD_COMPRESSION_SUM:
        return scalar_dev(b->S);
    case SPHX_FIELD_INTERPOLATED:
        return scalar_dev(b->interp);
    case SPHX_FIELD_FORCE:
        return soa_dev(b->force, T(0));
    case SPHX_FIELD_FORCE_PRIOR:
        return soa_dev(b->force_prior, T(0));
    case SPHX_FIELD_NORMAL:
        return soa_dev(b->normal, T(0));
    case SPHX_FIELD_WALL_VELOCITY:
        return soa_dev(b->wall_vel, T(0));
    case SPHX_FIELD_WALL_ACCELERATION:
        return soa_dev(b->wall_acc, T(0));
    case SPHX_FIELD_VELOCITY:
    {
        std::vector<Rec4<T>> tmp(n);
        if (write)
        {
            for (size_t i = 0; i < n; ++i)
            {
                tmp[i].x = T(h[i * D]);
                tmp[i].y = T(h[i * D + 1]);
                tmp[i].z = D == 3 ? T(h[i * D + 2]) : T(0);
                tmp[i].w = T(0);
            }
            SPHX_HIP(hipMemcpyAsync(b->vel, tmp.data(), n * sizeof(Rec4<T>), hipMemcpyHostToDevice, s));
            SPHX_HIP(hipStreamSynchronize(s));
        }
        else
        {
            SPHX_HIP(hipMemcpyAsync(tmp.data(), b->vel, n * sizeof(Rec4<T>), hipMemcpyDeviceToHost, s));
            SPHX_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i)
            {
                h[i * D] = (double)tmp[i].x;
                h[i * D + 1] = (double)tmp[i].y;
                if (D == 3)
                    h[i * D + 2] = (double)tmp[i].z;
            }
        }
        return SPHX_OK;
    }
    case SPHX_FIELD_CORRECTION_MATRIX:
    {
        // host: D x D row-major per particle; device: one array per component
        std::vector<T> tmp(n);
        for (int c = 0; c < D * D; ++c)
        {
            if (write)
            {
                for (size_t i = 0; i < n; ++i)
                    tmp[i] = T(h[i * D * D + c]);
                SPHX_HIP(hipMemcpyAsync(b->corr[c], tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
                SPHX_HIP(hipStreamSynchronize(s));
            }
            else
            {
                SPHX_HIP(hipMemcpyAsync(tmp.data(), b->corr[c], n * sizeof(T), hipMemcpyDeviceToHost, s));
                SPHX_HIP(hipStreamSynchronize(s));
                for (size_t i = 0; i < n; ++i)
                    h[i * D * D + c] = (double)tmp[i];
            }
        }
        return SPHX_OK;
    }
    case SPHX_FIELD_INDICATOR:
        return u32_dev((uint32_t *)b->indicator);
    case SPHX_FIELD_PREVIOUS_INDICATOR:
        return u32_dev((uint32_t *)b->prev_indicator);
    case SPHX_FIELD_POSITION_DIVERGENCE:
        return scalar_dev(b->pos_div);
    case SPHX_FIELD_VISCOUS_FORCE:
        return soa_dev(b->visc, T(0));
    case SPHX_FIELD_PREVIOUS_VISCOUS_FORCE:
        return soa_dev(b->prev_visc, T(0));
    case SPHX_FIELD_KERNEL_GRADIENT_INTEGRAL:
        return soa_dev(b->kgi, T(0));
    case SPHX_FIELD_ORIGINAL_ID:
        return u32_dev(b->original_id);
    case SPHX_FIELD_SORTED_ID:
        return u32_dev(b->sorted_id);
    case SPHX_FIELD_SORT_SEQUENCE:
        return u32_dev(b->sequence);
    default:
        return arg_error("unknown field id");
    }
}
} // namespace

} // namespace sphx

using namespace sphx;

namespace sphx
{
bool relation_ready(sphx_relation *r)
{
    return r && r->built && r->built_version == r->source->config_version &&
           r->target_version == r->target->config_version;
}
sphx_status check_fluid_relations(sphx_relation *inner, sphx_relation *const *walls, int n_wall, bool need_built)
{
    if (!inner || !inner->inner)
        return arg_error("an inner relation of the fluid body is required");
    if (n_wall < 0 || n_wall > 4 || (n_wall > 0 && !walls))
        return arg_error("0..4 contact relations are supported");
    if (inner->source->desc.kind != SPHX_BODY_FLUID)
        return arg_error("the inner relation's body is not a fluid body");
    if (need_built && !relation_ready(inner))
        return state_error("inner relation not updated since the last configuration change");
    for (int k = 0; k < n_wall; ++k)
    {
        sphx_relation *w = walls[k];
        if (!w || w->inner || w->source != inner->source)
            return arg_error("contact relations must start from the same fluid body");
        if (need_built && !relation_ready(w))
            return state_error("contact relation not updated since the last configuration change");
    }
    return SPHX_OK;
}
} // namespace sphx

namespace sphx
{
// dambreak.cpp:196-207: while (relaxation < advection dt) { dt = acoustic dt; 1st
// half; 2nd half; relaxation += dt }.  Each sub-step's dt, the loop condition
// and the time are evaluated on the device (k_substep_begin); launches past
// the end of the loop return at once.  The first batch has `batch` sub-steps;
// when it was too short the loop goes on one sub-step at a time.  On return
// the step state is in the context's pinned mirror (SLOT_STEP).
sphx_status advection_substeps(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                               const sphx_advection_config *cfg, uint32_t batch)
{
    sphx_body *b = inner->source;
    sphx_context *ctx = b->ctx;
    const bool corrected = cfg->correction == SPHX_CORRECTION_LINEAR;
    struct DeviceStep
    {
        sphx_context *c;
        explicit DeviceStep(sphx_context *c) : c(c) { c->device_step = true; }
        ~DeviceStep() { c->device_step = false; }
    } device_step(ctx);
    const unsigned long long *st = ctx->scalars_h + SLOT_STEP;
    sphx_status s;
    while (true)
    {
        for (uint32_t k = 0; k < batch; ++k)
        {
            if ((s = substep_begin(ctx, cfg->acoustic_cfl, b->h)))
                return s;
            // k_substep_begin read the signal max and reset the lanes for pass 2
            if (corrected)
                s = acoustic_1st_half_corrected(inner, walls, n_wall, cfg->riemann, 0.0, true);
            else
                s = acoustic_1st_half(inner, walls, n_wall, cfg->riemann, 0.0, true);
            if (s)
                return s;
            if (corrected)
                s = acoustic_2nd_half_corrected(inner, walls, n_wall, cfg->riemann, 0.0, false, true);
            else
                s = acoustic_2nd_half(inner, walls, n_wall, cfg->riemann, 0.0, false, true);
            if (s)
                return s;
        }
        if ((s = read_scalars(ctx, SLOT_STEP, STEP_SLOTS)))
            return s;
        if (st[STEP_ERROR])
        {
            set_error("non-finite acoustic time step: the fluid state diverged");
            return SPHX_E_NONFINITE;
        }
        double relaxation, adv_dt;
        std::memcpy(&relaxation, st + STEP_RELAXATION, sizeof(double));
        std::memcpy(&adv_dt, st + STEP_ADV_DT, sizeof(double));
        if (!(relaxation < adv_dt))
            return SPHX_OK;
        batch = 1;
    }
}

// the end of an advection step: position update, ParticleSortCK when due, the
// fluid cell linked list and the relations
sphx_status advection_tail(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                           sphx_cell_linked_list *cll, bool sort_now)
{
    sphx_body *b = inner->source;
    sphx_status s;
    if ((s = sphx_update_particle_position(b)))
        return s;
    if (sort_now && (s = sphx_particle_sort(b, cll)))
        return s;
    if ((s = sphx_update_cell_linked_list(cll)) || (s = sphx_update_relation(inner)))
        return s;
    for (int k = 0; k < n_wall; ++k)
        if ((s = sphx_update_relation(walls[k])))
            return s;
    return SPHX_OK;
}
} // namespace sphx

extern "C" {

const char *sphx_backend_name(void) { return "hip-gfx950"; }
const char *sphx_last_error(void) { return g_error.c_str(); }

sphx_status sphx_context_create(int device, sphx_context **out)
{
    if (!out)
        return arg_error("null out");
    int count = 0;
    SPHX_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return arg_error("device index out of range");
    SPHX_HIP(hipSetDevice(device));
    sphx_status st;
    // the workspace address of the allocation-free entry points, once per device
    if ((st = noalloc_scratch_init(device)))
        return st;
    std::unique_ptr<sphx_context> ctx(new sphx_context());
    ctx->device = device;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        ctx->max_lds = (size_t)lds;
    // on a failure below, the stream and the pinned mirror go back to the pools
    struct Release
    {
        sphx_context *c;
        ~Release()
        {
            if (!c)
                return;
            std::lock_guard<std::mutex> lock(g_reuse_mutex);
            if (c->stream)
                g_free_streams[c->device].push_back(c->stream);
            if (c->scalars_h)
                g_free_mirrors.push_back(c->scalars_h);
        }
    } release{ctx.get()};
    {
        std::lock_guard<std::mutex> lock(g_reuse_mutex);
        auto &streams = g_free_streams[device];
        if (!streams.empty())
        {
            ctx->stream = streams.back();
            streams.pop_back();
        }
        if (!g_free_mirrors.empty())
        {
            ctx->scalars_h = g_free_mirrors.back();
            g_free_mirrors.pop_back();
        }
    }
    if (!ctx->stream)
        SPHX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (Scratch *sc : {&ctx->sort_keys, &ctx->sort_vals, &ctx->sort_hist, &ctx->scan_blocks, &ctx->reduce_parts,
                        &ctx->gather_tmp, &ctx->cll_cursor})
        sc->stream = ctx->stream;
    if (!ctx->scalars_h)
        SPHX_HIP(hipHostMalloc(&ctx->scalars_h, kScalarSlots * sizeof(unsigned long long), hipHostMallocDefault));
    if ((st = dev_alloc(ctx->stream, (void **)&ctx->scalars_d, kScalarSlots * sizeof(unsigned long long))))
        return st;
    std::memset(ctx->scalars_h, 0, kScalarSlots * sizeof(unsigned long long));
    SPHX_HIP(hipMemcpyAsync(ctx->scalars_d, ctx->scalars_h, kScalarSlots * sizeof(unsigned long long),
                            hipMemcpyHostToDevice, ctx->stream));
    SPHX_HIP(hipStreamSynchronize(ctx->stream));
    release.c = nullptr;
    *out = ctx.release();
    return SPHX_OK;
}

sphx_status sphx_context_destroy(sphx_context *ctx)
{
    if (!ctx)
        return SPHX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drain_profile(ctx);
    for (hipEvent_t e : ctx->event_pool)
        (void)hipEventDestroy(e);
    for (Scratch *sc : {&ctx->sort_keys, &ctx->sort_vals, &ctx->sort_hist, &ctx->scan_blocks, &ctx->reduce_parts,
                        &ctx->gather_tmp, &ctx->cll_cursor})
        sc->release();
    dev_free(ctx->stream, ctx->scalars_d);
    (void)hipStreamSynchronize(ctx->stream);
    {
        std::lock_guard<std::mutex> lock(g_reuse_mutex);
        g_free_streams[ctx->device].push_back(ctx->stream);
        g_free_mirrors.push_back(ctx->scalars_h);
    }
    delete ctx;
    return SPHX_OK;
}

sphx_status sphx_context_synchronize(sphx_context *ctx)
{
    if (!ctx)
        return arg_error("null context");
    SPHX_HIP(hipStreamSynchronize(ctx->stream));
    return SPHX_OK;
}

sphx_status sphx_profile_enable(sphx_context *ctx, int enable)
{
    if (!ctx)
        return arg_error("null context");
    SPHX_HIP(hipSetDevice(ctx->device));
    SPHX_HIP(hipStreamSynchronize(ctx->stream));
    drain_profile(ctx);
    ctx->profile.clear();
    if (enable < 0 || enable > 2)
        return arg_error("profile enable is 0, 1 or 2");
    ctx->profiling = enable != 0;
    ctx->profile_passes_only = enable == 2;
    return SPHX_OK;
}

sphx_status sphx_profile_read(sphx_context *ctx, const char *stage, double *total_ms, uint64_t *launches)
{
    if (!ctx || !stage)
        return arg_error("null argument");
    SPHX_HIP(hipSetDevice(ctx->device));
    SPHX_HIP(hipStreamSynchronize(ctx->stream));
    drain_profile(ctx);
    auto it = ctx->profile.find(stage);
    if (total_ms)
        *total_ms = it == ctx->profile.end() ? 0.0 : it->second.first;
    if (launches)
        *launches = it == ctx->profile.end() ? 0 : it->second.second;
    return SPHX_OK;
}

namespace
{
// device blocks allocated by a create function, given back (dev_free) unless
// the function succeeds and calls keep()
struct AllocGuard
{
    hipStream_t stream;
    std::vector<void *> blocks;
    bool kept = false;
    void add(void *p)
    {
        if (p)
            blocks.push_back(p);
    }
    void keep() { kept = true; }
    ~AllocGuard()
    {
        if (!kept)
            for (void *p : blocks)
                sphx::dev_free(stream, p);
    }
};
} // namespace

sphx_status sphx_body_create(sphx_context *ctx, const sphx_body_desc *d, sphx_body **out)
{
    if (!ctx || !d || !out)
        return arg_error("null argument");
    if (d->dim != 2 && d->dim != 3)
        return arg_error("dim must be 2 or 3");
    if (d->precision != SPHX_FP64 && d->precision != SPHX_MIXED)
        return arg_error("precision must be SPHX_FP64 or SPHX_MIXED");
    if (!(d->spacing_ref > 0) || !(d->h_spacing_ratio > 0))
        return arg_error("spacing_ref and h_spacing_ratio must be positive");
    if (d->kind == SPHX_BODY_FLUID && (!(d->rho0 > 0) || !(d->c0 > 0)))
        return arg_error("fluid bodies need positive rho0 and c0");
    SPHX_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<sphx_body> b(new sphx_body());
    b->ctx = ctx;
    b->desc = *d;
    b->D = d->dim;
    b->P = d->precision;
    b->n = d->n_particles;
    // SPHAdaptation: h = h_spacing_ratio * (resolution / refinement), adaptation.cpp:12-19
    b->h = d->h_spacing_ratio * (d->spacing_ref / 1.0);
    const size_t n = b->n;
    const size_t T = b->P == P_FP64 ? sizeof(double) : sizeof(float);
    const size_t R4 = b->P == P_FP64 ? sizeof(Rec4<double>) : sizeof(Rec4<float>);
    b->tsize = T;
    // on a failure below, the blocks allocated so far go back to the cache
    struct Undo
    {
        sphx_body *b;
        ~Undo()
        {
            if (b)
                for (void *p : b->allocations)
                    dev_free(b->ctx->stream, p);
        }
    } undo{b.get()};
    sphx_status s;
    for (int k = 0; k < 3; ++k)
    {
        if ((s = alloc(b.get(), (void **)&b->pos[k], n * sizeof(double))) ||
            (s = alloc(b.get(), (void **)&b->dpos[k], n * sizeof(double))) ||
            (s = alloc(b.get(), &b->force[k], n * T)) || (s = alloc(b.get(), &b->force_prior[k], n * T)) ||
            (s = alloc(b.get(), &b->normal[k], n * T)) || (s = alloc(b.get(), &b->wall_vel[k], n * T)) ||
            (s = alloc(b.get(), &b->wall_acc[k], n * T)))
            return s;
    }
    void **scalars[] = {&b->Vol, &b->mass, &b->rho, &b->p, &b->C, &b->dC, &b->Vol_ref, &b->S, &b->interp};
    for (void **f : scalars)
        if ((s = alloc(b.get(), f, n * T)))
            return s;
    if (b->P == P_FP64)
    {
        if ((s = alloc(b.get(), (void **)&b->pos_f32, n * sizeof(Rec4<float>))))
            return s;
    }
    if ((s = alloc(b.get(), &b->posvol, n * R4)) || (s = alloc(b.get(), &b->vel, n * R4)) ||
        (s = alloc(b.get(), (void **)&b->original_id, n * 4)) || (s = alloc(b.get(), (void **)&b->sorted_id, n * 4)) ||
        (s = alloc(b.get(), (void **)&b->sequence, n * 4)) || (s = alloc(b.get(), (void **)&b->perm, n * 4)))
        return s;
    // correction matrices, their neighbour sum and the surface indicators: fluid bodies only
    const bool fluid = d->kind == SPHX_BODY_FLUID;
    const size_t nf = fluid ? n : 0;
    for (int c = 0; c < b->D * b->D; ++c)
        if ((s = alloc(b.get(), &b->corr[c], nf * T)))
            return s;
    for (int k = 0; k < b->D; ++k)
        if ((s = alloc(b.get(), &b->grad_sum[k], nf * T)))
            return s;
    if ((s = alloc(b.get(), (void **)&b->indicator, nf * 4)) ||
        (s = alloc(b.get(), (void **)&b->prev_indicator, nf * 4)) || (s = alloc(b.get(), &b->pos_div, nf * T)))
        return s;
    for (int k = 0; k < b->D; ++k)
        if ((s = alloc(b.get(), &b->visc[k], nf * T)) || (s = alloc(b.get(), &b->prev_visc[k], nf * T)) ||
            (s = alloc(b.get(), &b->kgi[k], nf * T)))
            return s;
    // zero every field, then the reference's initial values
    for (int k = 0; k < 3; ++k)
    {
        SPHX_HIP(hipMemsetAsync(b->pos[k], 0, n * sizeof(double), ctx->stream));
        SPHX_HIP(hipMemsetAsync(b->dpos[k], 0, n * sizeof(double), ctx->stream));
        for (void *f : {b->force[k], b->force_prior[k], b->normal[k], b->wall_vel[k], b->wall_acc[k]})
            SPHX_HIP(hipMemsetAsync(f, 0, n * T, ctx->stream));
    }
    for (void **f : scalars)
        SPHX_HIP(hipMemsetAsync(*f, 0, n * T, ctx->stream));
    SPHX_HIP(hipMemsetAsync(b->posvol, 0, n * R4, ctx->stream));
    if (b->P == P_MIXED)
        b->pos_f32 = (Rec4<float> *)b->posvol; // (x, y, z|Vol) in float already
    else
        SPHX_HIP(hipMemsetAsync(b->pos_f32, 0, n * sizeof(Rec4<float>), ctx->stream));
    SPHX_HIP(hipMemsetAsync(b->vel, 0, n * R4, ctx->stream));
    SPHX_HIP(hipMemsetAsync(b->sequence, 0, n * 4, ctx->stream));
    if (b->P == P_FP64)
    {
        if ((s = fill<double>(ctx, b->rho, n, d->rho0)) || (s = fill<double>(ctx, b->C, n, 1.0)) ||
            (s = fill<double>(ctx, b->S, n, 1.0)))
            return s;
    }
    else
    {
        // MIXED stores C - 1 (= 0 for C = 1)
        if ((s = fill<float>(ctx, b->rho, n, (float)d->rho0)) || (s = fill<float>(ctx, b->S, n, 1.0f)))
            return s;
    }
    if ((s = fill_iota(ctx, b->original_id, n)) || (s = fill_iota(ctx, b->sorted_id, n)))
        return s;
    // LinearCorrectionMatrix = identity (kernel_correction_ck.hpp:14-15); Indicator = 0,
    // PositionDivergence = 0, PreviousSurfaceIndicator = 1 (surface_indication_ck.hpp:17-18, 41)
    for (int c = 0; c < b->D * b->D; ++c)
    {
        const bool diag = c % (b->D + 1) == 0;
        if ((s = b->P == P_FP64 ? fill<double>(ctx, b->corr[c], nf, diag ? 1.0 : 0.0)
                                : fill<float>(ctx, b->corr[c], nf, diag ? 1.0f : 0.0f)))
            return s;
    }
    if ((s = fill<int32_t>(ctx, b->indicator, nf, 0)) || (s = fill<int32_t>(ctx, b->prev_indicator, nf, 1)))
        return s;
    if (nf)
    {
        SPHX_HIP(hipMemsetAsync(b->pos_div, 0, nf * T, ctx->stream));
        for (int k = 0; k < b->D; ++k)
            for (void *f : {b->grad_sum[k], b->visc[k], b->prev_visc[k], b->kgi[k]})
                SPHX_HIP(hipMemsetAsync(f, 0, nf * T, ctx->stream));
    }
    SPHX_HIP(hipStreamSynchronize(ctx->stream));
    undo.b = nullptr;
    *out = b.release();
    return SPHX_OK;
}

sphx_status sphx_body_destroy(sphx_body *b)
{
    if (!b)
        return SPHX_OK;
    for (void *p : b->allocations)
        dev_free(b->ctx->stream, p);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return SPHX_OK;
}

sphx_status sphx_body_smoothing_length(const sphx_body *b, double *h)
{
    if (!b || !h)
        return arg_error("null argument");
    *h = b->h;
    return SPHX_OK;
}

static sphx_status body_rw(sphx_body *b, int field, void *host, size_t n, bool write)
{
    if (!b || !host)
        return arg_error("null argument");
    if (n != b->n)
        return arg_error("particle count mismatch");
    if (field < 0 || field >= SPHX_FIELD_COUNT)
        return arg_error("unknown field id");
    if (field >= SPHX_FIELD_CORRECTION_MATRIX && b->desc.kind != SPHX_BODY_FLUID)
        return arg_error("correction, surface-indication, viscous and transport fields belong to fluid bodies");
    if (b->n == 0)
        return SPHX_OK;
    SPHX_HIP(hipSetDevice(b->ctx->device));
    sphx_status s = b->P == P_FP64 ? transfer<double>(b, field, host, write) : transfer<float>(b, field, host, write);
    if (s || !write)
        return s;
    if (field == SPHX_FIELD_CORRECTION_MATRIX)
        b->state_version++;
    if (field == SPHX_FIELD_VOLUME && !b->vol_written)
    {
        // first write: VolumetricMeasureRef = Vol (density_regularization.hpp:18-20),
        // Mass = rho0 * Vol (materials/base_material.cpp:37-40)
        b->vol_written = true;
        const double *v = (const double *)host;
        std::vector<double> m(b->n);
        for (size_t i = 0; i < b->n; ++i)
            m[i] = b->desc.rho0 * v[i];
        if ((s = body_rw(b, SPHX_FIELD_VOLUME_REF, host, n, true)) ||
            (s = body_rw(b, SPHX_FIELD_MASS, m.data(), n, true)))
            return s;
    }
    if (field == SPHX_FIELD_POSITION || field == SPHX_FIELD_VOLUME)
    {
        if (field == SPHX_FIELD_POSITION)
            b->config_version++;
        if ((s = refresh_posvol(b)))
            return s;
        SPHX_HIP(hipStreamSynchronize(b->ctx->stream));
    }
    return SPHX_OK;
}

sphx_status sphx_body_write(sphx_body *b, int field, const void *host, size_t n)
{
    return body_rw(b, field, const_cast<void *>(host), n, true);
}
sphx_status sphx_body_read(sphx_body *b, int field, void *host, size_t n)
{
    return body_rw(b, field, host, n, false);
}

sphx_status sphx_cell_linked_list_create(sphx_body *b, const double *lo, const double *up, double gs,
                                         uint32_t buffer_width, sphx_cell_linked_list **out)
{
    if (!b || !lo || !up || !out || !(gs > 0))
        return arg_error("bad cell linked list arguments");
    std::unique_ptr<sphx_cell_linked_list> c(new sphx_cell_linked_list());
    c->body = b;
    // Mesh(tentative_bounds, grid_spacing, buffer_width), base_mesh.cpp:6-16
    MeshParams &m = c->mesh;
    m.gs = gs;
    uint64_t ncells = 1;
    const double buffer = double(buffer_width) * gs;
    for (int k = 0; k < 3; ++k)
    {
        if (k < b->D)
        {
            m.lower[k] = lo[k] - buffer;
            double dim = up[k] + buffer - m.lower[k];
            double g = std::ceil(dim / gs);
            if (!(g >= 1.0) || g > 1.0e9)
                return arg_error("cell linked list bounds/grid spacing give an invalid mesh");
            m.all_grid_points[k] = (int)g + 1;
            m.all_cells[k] = m.all_grid_points[k] - 1;
            ncells *= (uint64_t)m.all_cells[k];
        }
        else
        {
            m.lower[k] = 0;
            m.all_grid_points[k] = 2;
            m.all_cells[k] = 1;
        }
    }
    if (ncells >= (1ull << 31))
        return arg_error("cell linked list has more than 2^31 cells");
    c->ncells = ncells;
    int bits = 0;
    while (bits < 32 && ((ncells - 1) >> bits) != 0)
        ++bits;
    c->key_bits = bits;
    SPHX_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    sphx_status st;
    AllocGuard guard{s};
    st = dev_alloc(s, (void **)&c->cell_offset, (ncells + 1) * sizeof(uint32_t));
    guard.add(c->cell_offset);
    if (!st)
        st = dev_alloc(s, (void **)&c->particle_index, std::max<size_t>(b->n, 1) * sizeof(uint32_t));
    guard.add(c->particle_index);
    if (!st)
        st = dev_alloc(s, (void **)&c->keys, std::max<size_t>(b->n, 1) * sizeof(uint32_t));
    guard.add(c->keys);
    if (st)
        return st;
    guard.keep();
    *out = c.release();
    return SPHX_OK;
}

sphx_status sphx_cell_linked_list_destroy(sphx_cell_linked_list *c)
{
    if (!c)
        return SPHX_OK;
    hipStream_t s = c->body->ctx->stream;
    for (void *p : {(void *)c->cell_offset, (void *)c->particle_index, (void *)c->keys, (void *)c->near})
        dev_free(s, p);
    (void)hipStreamSynchronize(s);
    delete c;
    return SPHX_OK;
}

sphx_status sphx_cell_linked_list_mesh(const sphx_cell_linked_list *c, int *all_cells, double *lower, double *gs)
{
    if (!c)
        return arg_error("null cell linked list");
    for (int k = 0; k < c->body->D; ++k)
    {
        if (all_cells)
            all_cells[k] = c->mesh.all_cells[k];
        if (lower)
            lower[k] = c->mesh.lower[k];
    }
    if (gs)
        *gs = c->mesh.gs;
    return SPHX_OK;
}

sphx_status sphx_update_cell_linked_list(sphx_cell_linked_list *c)
{
    if (!c)
        return arg_error("null cell linked list");
    SPHX_HIP(hipSetDevice(c->body->ctx->device));
    sphx_status s = cell_linked_list_update(c);
    if (s)
        return s;
    c->built_version = c->body->config_version;
    return SPHX_OK;
}

sphx_status sphx_update_cell_linked_list_noalloc(sphx_cell_linked_list *c)
{
    if (!c)
        return arg_error("null cell linked list");
    SPHX_HIP(hipSetDevice(c->body->ctx->device));
    sphx_status s = cell_linked_list_update_noalloc(c);
    if (s)
        return s;
    c->built_version = c->body->config_version;
    return SPHX_OK;
}

sphx_status sphx_cell_linked_list_read(sphx_cell_linked_list *c, uint32_t *cell_offset, size_t n_off,
                                       uint32_t *particle_index, size_t n_idx)
{
    if (!c)
        return arg_error("null cell linked list");
    if (!c->built || c->built_version != c->body->config_version)
        return state_error("cell linked list not updated since the last configuration change");
    hipStream_t s = c->body->ctx->stream;
    if (cell_offset)
    {
        if (n_off != c->ncells + 1)
        {
            set_error("cell_offset size must be n_cells + 1");
            return SPHX_E_CAPACITY;
        }
        SPHX_HIP(hipMemcpyAsync(cell_offset, c->cell_offset, n_off * 4, hipMemcpyDeviceToHost, s));
    }
    if (particle_index)
    {
        if (n_idx != c->body->n)
        {
            set_error("particle_index size must be n_particles");
            return SPHX_E_CAPACITY;
        }
        if (n_idx)
            SPHX_HIP(hipMemcpyAsync(particle_index, c->particle_index, n_idx * 4, hipMemcpyDeviceToHost, s));
    }
    SPHX_HIP(hipStreamSynchronize(s));
    return SPHX_OK;
}

static sphx_status relation_create(sphx_body *src, sphx_body *tar, sphx_cell_linked_list *cll, bool inner,
                                   sphx_relation **out)
{
    std::unique_ptr<sphx_relation> r(new sphx_relation());
    r->inner = inner;
    r->source = src;
    r->target = tar;
    r->cll = cll;
    // Neighbor<SPHAdaptation,SPHAdaptation>: src_inv_h = 1/h_src, inv_h = 1/max(h_src, h_tar)
    r->src_inv_h = 1.0 / src->h;
    r->inv_h = 1.0 / std::max(src->h, tar->h);
    SPHX_HIP(hipSetDevice(src->ctx->device));
    hipStream_t stream = src->ctx->stream;
    // on a failure below, the blocks allocated so far go back to the cache
    struct Undo
    {
        sphx_relation *r;
        hipStream_t s;
        ~Undo()
        {
            if (!r)
                return;
            for (void *p : {(void *)r->neighbor_size, (void *)r->particle_offset, (void *)r->chunk_offset,
                            (void *)r->stage_count, (void *)r->stage_keys, r->poly})
                if (p)
                    dev_free(s, p);
        }
    } undo{r.get(), stream};
    {
        sphx_status st;
        if ((st = dev_alloc(stream, (void **)&r->neighbor_size, (size_t)(src->n + 1) * sizeof(uint32_t))) ||
            (st = dev_alloc(stream, (void **)&r->particle_offset, (size_t)(src->n + 1) * sizeof(uint32_t))) ||
            (st = dev_alloc(stream, (void **)&r->chunk_offset, (size_t)((src->n + 63) / 64 + 1) * sizeof(uint32_t))))
            return st;
    }
    SPHX_HIP(hipMemsetAsync(r->neighbor_size, 0, (size_t)(src->n + 1) * sizeof(uint32_t), src->ctx->stream));
    if (inner && src->desc.kind == SPHX_BODY_FLUID && src->n > 0)
    {
        // LDS staging of the acoustic passes: the 2nd pass stages two 4-wide
        // records per distinct neighbour; keep 8 KiB for the static tables
        // (MIXED only: FP64 records would halve the occupancy of the passes)
        const size_t rec2 = 2 * (src->P == P_FP64 ? sizeof(Rec4<double>) : sizeof(Rec4<float>));
        const size_t cap = kStageCap;
        if (src->P == P_MIXED && cap * rec2 + 8192 <= std::min<size_t>(src->ctx->max_lds, 65536))
        {
            const size_t nblocks = (src->n + kStageBlock - 1) / kStageBlock;
            r->staged = true;
            r->stage_cap = (uint32_t)cap;
            sphx_status st;
            if ((st = dev_alloc(stream, (void **)&r->stage_count, nblocks * sizeof(uint32_t))) ||
                (st = dev_alloc(stream, (void **)&r->stage_keys, nblocks * cap * sizeof(uint32_t))))
                return st;
        }
    }
    sphx_status s = build_kernel_poly(r.get());
    if (s)
        return s;
    undo.r = nullptr;
    *out = r.release();
    return SPHX_OK;
}

sphx_status sphx_relation_create_inner(sphx_body *b, sphx_cell_linked_list *cll, sphx_relation **out)
{
    if (!b || !cll || !out || cll->body != b)
        return arg_error("inner relation needs the body's own cell linked list");
    return relation_create(b, b, cll, true, out);
}

sphx_status sphx_relation_create_contact(sphx_body *src, sphx_body *tar, sphx_cell_linked_list *cll,
                                         sphx_relation **out)
{
    if (!src || !tar || !cll || !out || cll->body != tar)
        return arg_error("contact relation needs the target body's cell linked list");
    if (src->D != tar->D || src->P != tar->P || src->ctx != tar->ctx)
        return arg_error("contact relation bodies differ in dimension, precision or context");
    return relation_create(src, tar, cll, false, out);
}

sphx_status sphx_relation_destroy(sphx_relation *r)
{
    if (!r)
        return SPHX_OK;
    hipStream_t s = r->source->ctx->stream;
    for (void *p : {(void *)r->neighbor_size, (void *)r->particle_offset, (void *)r->neighbor_index,
                    (void *)r->chunk_offset, (void *)r->slice, (void *)r->stage_count, (void *)r->stage_keys,
                    (void *)r->local_slice, (void *)r->poly})
        dev_free(s, p);
    (void)hipStreamSynchronize(s);
    delete r;
    return SPHX_OK;
}

sphx_status sphx_update_relation(sphx_relation *r)
{
    if (!r)
        return arg_error("null relation");
    sphx_cell_linked_list *c = r->cll;
    if (!c->built || c->built_version != c->body->config_version)
        return state_error("target cell linked list not updated since the last configuration change");
    sphx_context *ctx = r->source->ctx;
    SPHX_HIP(hipSetDevice(ctx->device));
    sphx_status s;
    // one scan builds the slices (and the staging sets); grow the slice width
    // and rebuild when a particle has more neighbours than it, as the
    // reference grows NeighborIndex on overflow (update_body_relation.hpp:145-155)
    const size_t n = r->source->n;
    const size_t nchunks = (n + 63) / 64;
    uint32_t kmax = r->kmax ? r->kmax : (r->source->D == 3 ? 96u : 32u);
    if (!r->inner && (s = cell_near_flags(c)))
        return s;
    bool scan = false, done = false;
    for (int attempt = 0; attempt < 5 && !done; ++attempt)
    {
        const size_t need = std::max<size_t>(nchunks * kmax * 64, 64);
        if (need >= (1ull << 32))
        {
            set_error("neighbour slices exceed 2^32 entries");
            return SPHX_E_CAPACITY;
        }
        if (need > r->slice_capacity || !r->slice)
        {
            // released after the kernels already queued on the stream
            dev_free(ctx->stream, r->slice);
            r->slice = nullptr;
            dev_free(ctx->stream, r->local_slice);
            r->local_slice = nullptr;
            r->slice_capacity = 0;
            if ((s = dev_alloc(ctx->stream, (void **)&r->slice, need * sizeof(uint32_t))))
                return s;
            if (r->staged && (s = dev_alloc(ctx->stream, (void **)&r->local_slice, need * sizeof(uint16_t))))
                return s;
            r->slice_capacity = need;
        }
        if ((s = fill_u64(ctx, ctx->scalars_d + SLOT_NEEDED, 1, 0ull)) ||
            (s = relation_build(r, kmax, ctx->scalars_d + SLOT_NEEDED, scan)) ||
            (s = read_scalars(ctx, SLOT_NEEDED, 1)))
            return s;
        const unsigned long long needed = ctx->scalars_h[SLOT_NEEDED];
        if (needed <= kmax)
        {
            done = true;
            break;
        }
        if (needed >= kTooDense)
        {
            scan = true; // a cell row too dense for the tiled build's row code
            continue;
        }
        if (needed > 60000)
        {
            set_error("a particle has more than 60000 neighbours: positions collapsed");
            return SPHX_E_CAPACITY;
        }
        kmax = (uint32_t)((needed + needed / 4 + 7) / 8 * 8);
    }
    if (!done)
    {
        set_error("neighbour lists did not fit after growing the slice width");
        return SPHX_E_CAPACITY;
    }
    r->kmax = kmax;
    r->slice_total = nchunks * kmax * 64;
    if ((s = relation_stage(r)))
        return s;
    r->stage_stale = false;
    r->csr_valid = false; // the CSR view (offsets and indices) is built when read
    r->built = true;
    r->builds++;
    r->built_version = r->source->config_version;
    r->target_version = r->target->config_version;
    return SPHX_OK;
}

sphx_status sphx_relation_read(sphx_relation *r, uint32_t *size, uint32_t *offset, uint32_t *index, size_t capacity,
                               size_t *total)
{
    if (!r)
        return arg_error("null relation");
    if (!relation_ready(r))
        return state_error("relation not updated since the last configuration change");
    sphx_context *ctx = r->source->ctx;
    hipStream_t s = ctx->stream;
    SPHX_HIP(hipSetDevice(ctx->device));
    size_t n = r->source->n;
    if (!r->csr_valid)
    {
        // CSR offsets of the reference layout (NeighborSize scan), then the indices
        sphx_status st = exclusive_scan_u32(ctx, r->neighbor_size, r->particle_offset, n + 1, nullptr);
        if (st)
            return st;
        uint32_t t = 0;
        SPHX_HIP(hipMemcpyAsync(&t, r->particle_offset + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        SPHX_HIP(hipStreamSynchronize(s));
        r->total = t;
        if (t > r->capacity || !r->neighbor_index)
        {
            dev_free(s, r->neighbor_index);
            r->neighbor_index = nullptr;
            r->capacity = 0;
            const size_t cap = std::max<size_t>(t, 64);
            if ((st = dev_alloc(s, (void **)&r->neighbor_index, cap * sizeof(uint32_t))))
                return st;
            r->capacity = cap;
        }
        if ((st = relation_to_csr(r)))
            return st;
        r->csr_valid = true;
    }
    if (total)
        *total = r->total;
    if (size && n)
        SPHX_HIP(hipMemcpyAsync(size, r->neighbor_size, n * 4, hipMemcpyDeviceToHost, s));
    if (offset)
        SPHX_HIP(hipMemcpyAsync(offset, r->particle_offset, (n + 1) * 4, hipMemcpyDeviceToHost, s));
    if (index)
    {
        if (capacity < r->total)
        {
            set_error("neighbor_index capacity too small");
            return SPHX_E_CAPACITY;
        }
        if (r->total)
            SPHX_HIP(hipMemcpyAsync(index, r->neighbor_index, r->total * 4, hipMemcpyDeviceToHost, s));
    }
    SPHX_HIP(hipStreamSynchronize(s));
    return SPHX_OK;
}

sphx_status sphx_particle_sort(sphx_body *b, sphx_cell_linked_list *c)
{
    if (!b || !c || c->body != b)
        return arg_error("sort needs the body's own cell linked list");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    ProfileScope prof(b->ctx, "particle_sort");
    sphx_status s;
    if ((s = morton_sequence(b, c)))
        return s;
    // UpdateSortableVariables over the evolving variables only
    for (int k = 0; k < b->D; ++k)
        if ((s = gather_double(b, &b->pos[k], b->perm)) || (s = gather_real(b, &b->force_prior[k], b->perm)))
            return s;
    if ((s = gather_real(b, &b->Vol, b->perm)) || (s = gather_u32(b->ctx, b->original_id, b->perm, b->n)) ||
        (s = gather_rec4(b, &b->vel, b->perm)) || (s = gather_real(b, &b->mass, b->perm)) ||
        (s = gather_real(b, &b->C, b->perm)) || (s = gather_real(b, &b->dC, b->perm)) ||
        (s = gather_real(b, &b->Vol_ref, b->perm)) || (s = sorted_id_update(b)) || (s = refresh_posvol(b)))
        return s;
    // evolving once a free-surface indication exists (surface_indication_ck.hpp:43)
    if (b->desc.kind == SPHX_BODY_FLUID && (s = gather_u32(b->ctx, (uint32_t *)b->prev_indicator, b->perm, b->n)))
        return s;
    // evolving with a viscous force (force_prior_ck.cpp:14)
    for (int k = 0; b->desc.kind == SPHX_BODY_FLUID && k < b->D; ++k)
        if ((s = gather_real(b, &b->prev_visc[k], b->perm)))
            return s;
    b->config_version++;
    return SPHX_OK;
}

sphx_status sphx_gravity_force(sphx_body *b, const double *g)
{
    if (!b || !g)
        return arg_error("null argument");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    return gravity_force(b, g);
}

sphx_status sphx_density_regularization(sphx_body *b, int flow)
{
    if (!b || (flow != SPHX_FLOW_INTERNAL && flow != SPHX_FLOW_FREE_SURFACE))
        return arg_error("bad density regularization arguments");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    return density_regularization(b, flow);
}

sphx_status sphx_advection_step_setup(sphx_body *b)
{
    if (!b)
        return arg_error("null body");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    return advection_step_setup(b);
}

sphx_status sphx_update_particle_position(sphx_body *b)
{
    if (!b)
        return arg_error("null body");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    sphx_status s = update_position(b);
    if (!s)
        b->config_version++;
    return s;
}

sphx_status sphx_compression_summation(sphx_relation *inner, sphx_relation *const *contact, int n)
{
    sphx_status s = check_fluid_relations(inner, contact, n, true);
    if (s)
        return s;
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return compression_summation(inner, contact, n);
}

sphx_status sphx_acoustic_step_1st_half(sphx_relation *inner, sphx_relation *const *walls, int n, int riemann,
                                        double dt)
{
    sphx_status s = check_fluid_relations(inner, walls, n, true);
    if (s)
        return s;
    if (riemann < SPHX_RIEMANN_ACOUSTIC || riemann > SPHX_RIEMANN_DISSIPATIVE)
        return arg_error("unknown Riemann solver");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return acoustic_1st_half(inner, walls, n, riemann, dt, false);
}

sphx_status sphx_acoustic_step_2nd_half(sphx_relation *inner, sphx_relation *const *walls, int n, int riemann,
                                        double dt)
{
    sphx_status s = check_fluid_relations(inner, walls, n, true);
    if (s)
        return s;
    if (riemann < SPHX_RIEMANN_ACOUSTIC || riemann > SPHX_RIEMANN_DISSIPATIVE)
        return arg_error("unknown Riemann solver");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return acoustic_2nd_half(inner, walls, n, riemann, dt, true, false);
}

sphx_status sphx_acoustic_step_2nd_half_traversal(sphx_relation *inner, sphx_relation *const *walls, int n,
                                                  int riemann, double dt)
{
    sphx_status s = check_fluid_relations(inner, walls, n, true);
    if (s)
        return s;
    if (riemann < SPHX_RIEMANN_ACOUSTIC || riemann > SPHX_RIEMANN_DISSIPATIVE)
        return arg_error("unknown Riemann solver");
    const sphx_cell_linked_list *c = inner->cll;
    if (!c->built || c->built_version != c->body->config_version)
        return state_error("cell linked list not updated since the last configuration change");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return acoustic_2nd_half_traversal(inner, walls, n, riemann, dt);
}

sphx_status sphx_acoustic_step_1st_half_corrected(sphx_relation *inner, sphx_relation *const *walls, int n,
                                                  int riemann, double dt)
{
    sphx_status s = check_fluid_relations(inner, walls, n, true);
    if (s)
        return s;
    if (riemann < SPHX_RIEMANN_ACOUSTIC || riemann > SPHX_RIEMANN_DISSIPATIVE)
        return arg_error("unknown Riemann solver");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return acoustic_1st_half_corrected(inner, walls, n, riemann, dt, false);
}

sphx_status sphx_acoustic_step_2nd_half_corrected(sphx_relation *inner, sphx_relation *const *walls, int n,
                                                  int riemann, double dt)
{
    sphx_status s = check_fluid_relations(inner, walls, n, true);
    if (s)
        return s;
    if (riemann < SPHX_RIEMANN_ACOUSTIC || riemann > SPHX_RIEMANN_DISSIPATIVE)
        return arg_error("unknown Riemann solver");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return acoustic_2nd_half_corrected(inner, walls, n, riemann, dt, true, false);
}

sphx_status sphx_linear_correction_matrix(sphx_relation *inner, sphx_relation *const *contact, int n, double alpha)
{
    sphx_status s = check_fluid_relations(inner, contact, n, true);
    if (s)
        return s;
    if (!std::isfinite(alpha))
        return arg_error("alpha must be finite");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return linear_correction_matrix(inner, contact, n, alpha);
}

sphx_status sphx_free_surface_indication(sphx_relation *inner, sphx_relation *const *contact, int n)
{
    sphx_status s = check_fluid_relations(inner, contact, n, true);
    if (s)
        return s;
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return free_surface_indication(inner, contact, n);
}

sphx_status sphx_viscous_force(sphx_relation *inner, sphx_relation *const *walls, int n, double mu, int correction)
{
    sphx_status s = check_fluid_relations(inner, walls, n, true);
    if (s)
        return s;
    if (!(mu >= 0.0) || !std::isfinite(mu))
        return arg_error("viscosity must be finite and non-negative");
    if (correction != SPHX_CORRECTION_NONE && correction != SPHX_CORRECTION_LINEAR)
        return arg_error("unknown kernel correction");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return viscous_force(inner, walls, n, mu, correction == SPHX_CORRECTION_LINEAR);
}

sphx_status sphx_kernel_gradient_integral(sphx_relation *inner, sphx_relation *const *contact, int n, int correction)
{
    sphx_status s = check_fluid_relations(inner, contact, n, true);
    if (s)
        return s;
    if (correction != SPHX_CORRECTION_NONE && correction != SPHX_CORRECTION_LINEAR)
        return arg_error("unknown kernel correction");
    SPHX_HIP(hipSetDevice(inner->source->ctx->device));
    return kernel_gradient_integral(inner, contact, n, correction == SPHX_CORRECTION_LINEAR);
}

sphx_status sphx_transport_velocity_correction(sphx_body *b, double coefficient, int limiter, int bulk_only)
{
    if (!b || b->desc.kind != SPHX_BODY_FLUID)
        return arg_error("transport-velocity correction needs a fluid body");
    if (limiter != SPHX_LIMITER_NONE && limiter != SPHX_LIMITER_TRUNCATED_LINEAR)
        return arg_error("unknown limiter");
    if (!std::isfinite(coefficient))
        return arg_error("coefficient must be finite");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    return transport_velocity_correction(b, coefficient, limiter, bulk_only != 0);
}

// the max lanes: reset before a max reduction, combined when read
static sphx_status reset_max(sphx_context *ctx)
{
    return fill_u64(ctx, ctx->scalars_d + SLOT_MAX, (size_t)kMaxLanes * kMaxStride, kOrderedLowest);
}
static sphx_status read_max(sphx_context *ctx, double *out)
{
    sphx_status s = read_scalars(ctx, SLOT_MAX, kMaxLanes * kMaxStride);
    if (s)
        return s;
    unsigned long long m = 0;
    for (int k = 0; k < kMaxLanes; ++k)
        m = std::max(m, ctx->scalars_h[SLOT_MAX + k * kMaxStride]);
    *out = from_ordered_bits(m);
    return SPHX_OK;
}

// the max reductions of the time steps take +inf for a particle with a
// non-finite velocity, force or compression, which makes the step zero
static sphx_status finite_step(double dt, const char *which)
{
    if (std::isfinite(dt) && dt > 0.0)
        return SPHX_OK;
    set_error(std::string("non-finite ") + which + " time step: a particle's velocity, force or compression "
              "is not finite");
    return SPHX_E_NONFINITE;
}

sphx_status sphx_acoustic_time_step(sphx_body *b, double cfl, double *dt)
{
    if (!b || !dt)
        return arg_error("null argument");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    sphx_context *ctx = b->ctx;
    sphx_status s;
    if ((s = reset_max(ctx)) ||
        (s = acoustic_signal_max(b, ctx->scalars_d + SLOT_MAX)))
        return s;
    double m = 0;
    if ((s = read_max(ctx, &m)))
        return s;
    // AcousticTimeStepCK::FinishDynamics::Result, fluid_time_step_ck.hpp:30-36
    *dt = cfl * b->h / (m + 2.71051e-20);
    return finite_step(*dt, "acoustic");
}

sphx_status sphx_advection_time_step(sphx_body *b, double u_ref, double cfl, double *dt)
{
    if (!b || !dt)
        return arg_error("null argument");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    sphx_context *ctx = b->ctx;
    sphx_status s;
    if ((s = reset_max(ctx)) ||
        (s = advection_speed_max(b, ctx->scalars_d + SLOT_MAX)))
        return s;
    double m = 0;
    if ((s = read_max(ctx, &m)))
        return s;
    // AdvectionTimeStepCK::FinishDynamics::Result, fluid_time_step_ck.cpp:23-26
    *dt = cfl * b->h / (std::max(std::sqrt(m), u_ref) + 2.71051e-20);
    return finite_step(*dt, "advection");
}

sphx_status sphx_total_mechanical_energy(sphx_body *b, const double *g, double *e)
{
    if (!b || !g || !e)
        return arg_error("null argument");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    sphx_context *ctx = b->ctx;
    sphx_status s = mechanical_energy(b, g, (double *)(ctx->scalars_d + SLOT_SUM_BASE));
    if (s || (s = read_scalars(ctx, SLOT_SUM_BASE, 1)))
        return s;
    std::memcpy(e, ctx->scalars_h + SLOT_SUM_BASE, sizeof(double));
    return SPHX_OK;
}

sphx_status sphx_upper_front_in_axis(sphx_body *b, int axis, double *front)
{
    if (!b || !front || axis < 0 || axis >= b->D)
        return arg_error("bad upper front arguments");
    SPHX_HIP(hipSetDevice(b->ctx->device));
    sphx_context *ctx = b->ctx;
    sphx_status s;
    if ((s = reset_max(ctx)) ||
        (s = axis_max(b, axis, ctx->scalars_d + SLOT_MAX)))
        return s;
    return read_max(ctx, front);
}

sphx_status sphx_interpolate_contact(sphx_relation *r, int field, double *out, size_t n_obs)
{
    if (!r || r->inner || !out)
        return arg_error("interpolation needs a contact relation");
    if (!relation_ready(r))
        return state_error("observer relation not updated since the last configuration change");
    if (n_obs != r->source->n)
        return arg_error("observer count mismatch");
    if (field != SPHX_FIELD_PRESSURE && field != SPHX_FIELD_DENSITY && field != SPHX_FIELD_COMPRESSION &&
        field != SPHX_FIELD_VELOCITY)
        return arg_error("unsupported interpolation field");
    SPHX_HIP(hipSetDevice(r->source->ctx->device));
    sphx_status s;
    if (field == SPHX_FIELD_VELOCITY)
    {
        // Interpolation<Contact<Vecd>>, one component at a time
        const int D = r->source->D;
        std::vector<double> comp(n_obs);
        for (int k = 0; k < D; ++k)
        {
            if ((s = interpolate_contact(r, field, k)) ||
                (s = sphx_body_read(r->source, SPHX_FIELD_INTERPOLATED, comp.data(), n_obs)))
                return s;
            for (size_t i = 0; i < n_obs; ++i)
                out[i * D + k] = comp[i];
        }
        return SPHX_OK;
    }
    if ((s = interpolate_contact(r, field, 0)))
        return s;
    if ((s = sphx_body_read(r->source, SPHX_FIELD_INTERPOLATED, out, n_obs)))
        return s;
    if (field == SPHX_FIELD_COMPRESSION && r->source->P == P_MIXED)
        for (size_t i = 0; i < n_obs; ++i)
            out[i] += 1.0; // MIXED stores C - 1
    return SPHX_OK;
}

sphx_status sphx_advection_step(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                                 sphx_cell_linked_list *cll, const sphx_advection_config *cfg, double *time,
                                 uint32_t *steps)
{
    sphx_status s = check_fluid_relations(inner, walls, n_wall, true);
    if (s)
        return s;
    if (!cfg || !time || !cll || cll != inner->cll)
        return arg_error("advection step needs the fluid cell linked list, a config and a time pointer");
    if (cfg->correction != SPHX_CORRECTION_NONE && cfg->correction != SPHX_CORRECTION_LINEAR)
        return arg_error("unknown kernel correction");
    sphx_body *b = inner->source;
    sphx_context *ctx = b->ctx;
    SPHX_HIP(hipSetDevice(ctx->device));
    if ((s = compression_summation(inner, walls, n_wall)) || (s = density_regularization(b, cfg->flow_type)) ||
        (s = advection_step_setup(b)))
        return s;
    // AdvectionTimeStepCK on the device: the loop below runs without host
    // round trips, the host reads the loop state once per batch of sub-steps
    if ((s = reset_max(ctx)) || (s = advection_speed_max(b, ctx->scalars_d + SLOT_MAX)) ||
        (s = step_begin(ctx, cfg->u_ref, cfg->advection_cfl, b->h, *time)))
        return s;
    // dambreak.cpp:192-193
    if (cfg->surface_indication && (s = free_surface_indication(inner, walls, n_wall)))
        return s;
    const bool corrected = cfg->correction == SPHX_CORRECTION_LINEAR;
    if (corrected && (s = linear_correction_matrix(inner, walls, n_wall, cfg->correction_alpha)))
        return s;
    // AcousticTimeStepCK of the first sub-step from the current state (k_step_begin
    // left the max lanes reset)
    if ((s = acoustic_signal_max(b, ctx->scalars_d + SLOT_MAX)))
        return s;
    if ((s = advection_substeps(inner, walls, n_wall, cfg, ctx->last_substeps ? ctx->last_substeps : 1u)))
        return s;
    const unsigned long long *st = ctx->scalars_h + SLOT_STEP;
    const uint32_t count = (uint32_t)st[STEP_COUNT];
    std::memcpy(time, st + STEP_TIME, sizeof(double));
    ctx->last_substeps = count;
    if ((s = advection_tail(inner, walls, n_wall, cll, cfg->sort_now != 0)))
        return s;
    if (steps)
        *steps = count;
    return SPHX_OK;
}

} // extern "C"
