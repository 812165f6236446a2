// config_kernels.hip -- configuration dynamics on gfx950: exclusive scan, LDS
// radix sort, cell linked list, neighbour relations, particle sort gathers.
//
// Everything here is integer/index work whose results must be bit-exact with
// the CPU path on identical positions, so floating-point contraction is off
// for this translation unit: cell indices and the neighbour criterion are
// evaluated with the reference's separate IEEE operations.
#pragma clang fp contract(off)

#include "sphx_internal.hpp"

namespace sphx
{
namespace
{
constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems; // 2048 keys per workgroup
constexpr int kRadixBits = 4;
constexpr int kBuckets = 1 << kRadixBits;

inline unsigned grid_for(size_t n, size_t per) { return (unsigned)((n + per - 1) / per); }

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off)
            v += t;
    }
    return v;
}

// exclusive scan of one value per thread over a 256-thread block; returns the
// block total through *total (all threads)
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t wave_tot[kThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63)
        wave_tot[wave] = inc;
    __syncthreads();
    uint32_t base = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w)
    {
        uint32_t t = wave_tot[w];
        base += (w < wave) ? t : 0u;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return base + inc - v;
}

//---------------------------------------------------------------------------
// exclusive scan: tile reduce -> scan of tile sums -> tile scan
//---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t *in, size_t n, uint32_t *sums)
{
    size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n)
            s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = tot;
}

// single workgroup: exclusive scan of the tile sums in place; grand total to *total
__global__ void __launch_bounds__(kThreads) k_scan_sums(uint32_t *sums, uint32_t nsums,
                                                        unsigned long long *total)
{
    uint32_t carry = 0;
    for (uint32_t start = 0; start < nsums; start += kTile)
    {
        uint32_t base = start + threadIdx.x * kItems;
        uint32_t v[kItems];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kItems; ++k)
        {
            v[k] = (base + k < nsums) ? sums[base + k] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(s, &tot);
        uint32_t run = carry + ex;
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (base + k < nsums)
            {
                sums[base + k] = run;
                run += v[k];
            }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total)
        *total = carry;
}

__global__ void __launch_bounds__(kThreads) k_scan_apply(const uint32_t *in, uint32_t *out, size_t n,
                                                         const uint32_t *sums)
{
    size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + block_exclusive_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n)
        {
            out[base + k] = run;
            run += v[k];
        }
}

//---------------------------------------------------------------------------
// LSD radix sort of (key, value) pairs, 4-bit digits, stable.
//---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_radix_hist(const uint32_t *keys, size_t n, int shift,
                                                         uint32_t *hist, uint32_t ntiles)
{
    __shared__ uint32_t cnt[kBuckets];
    if (threadIdx.x < kBuckets)
        cnt[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n)
            atomicAdd(&cnt[(keys[base + k] >> shift) & (kBuckets - 1)], 1u);
    __syncthreads();
    if (threadIdx.x < kBuckets)
        hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(kThreads) k_radix_scatter(const uint32_t *keys_in, const uint32_t *vals_in,
                                                            uint32_t *keys_out, uint32_t *vals_out, size_t n,
                                                            int shift, const uint32_t *offsets, uint32_t ntiles)
{
    // cnt[d * 256 + t]: items of digit d held by thread t, scanned in (d, t) order
    __shared__ uint32_t cnt[kBuckets * kThreads];
    const int t = threadIdx.x;
#pragma unroll
    for (int d = 0; d < kBuckets; ++d)
        cnt[d * kThreads + t] = 0;
    size_t base = (size_t)blockIdx.x * kTile + (size_t)t * kItems;
    uint32_t key[kItems], val[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        if (base + k < n)
        {
            key[k] = keys_in[base + k];
            val[k] = vals_in[base + k];
            cnt[((key[k] >> shift) & (kBuckets - 1)) * kThreads + t] += 1;
        }
    }
    __syncthreads();
    // scan of the 4096 counters: thread t owns entries [16t, 16t+16)
    constexpr int per = kBuckets * kThreads / kThreads; // 16
    uint32_t local[per];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < per; ++k)
    {
        local[k] = cnt[t * per + k];
        s += local[k];
    }
    __syncthreads();
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < per; ++k)
    {
        cnt[t * per + k] = run;
        run += local[k];
    }
    __syncthreads();
    __shared__ uint32_t dstart[kBuckets];
    if (t < kBuckets)
        dstart[t] = cnt[t * kThreads];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        if (base + k < n)
        {
            uint32_t d = (key[k] >> shift) & (kBuckets - 1);
            uint32_t r = cnt[d * kThreads + t];
            cnt[d * kThreads + t] = r + 1; // column t is private to thread t
            size_t dst = (size_t)offsets[(size_t)d * ntiles + blockIdx.x] + (r - dstart[d]);
            keys_out[dst] = key[k];
            vals_out[dst] = val[k];
        }
    }
}

//---------------------------------------------------------------------------
// cell linked list: Mesh::LinearCellIndexFromPosition, base_mesh.hxx:9-26,73-78
//---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void cell_index(const MeshParams &m, const double *x, int *c)
{
#pragma unroll
    for (int k = 0; k < D; ++k)
    {
        int ci = (int)floor((x[k] - m.lower[k]) / m.gs);
        c[k] = min(max(ci, 0), m.all_grid_points[k] - 2);
    }
}
template <int D>
__device__ __forceinline__ uint32_t linear_cell(const MeshParams &m, const int *c)
{
    if (D == 2)
        return (uint32_t)(c[0] * m.all_cells[1] + c[1]);
    return (uint32_t)(c[0] * m.all_cells[1] * m.all_cells[2] + c[1] * m.all_cells[2] + c[2]);
}
// Mesh::MortonCode, base_mesh.hxx:91-99
__device__ __forceinline__ uint32_t morton_code(uint32_t x)
{
    x &= 0x3ff;
    x = (x | x << 16) & 0x30000ff;
    x = (x | x << 8) & 0x300f00f;
    x = (x | x << 4) & 0x30c30c3;
    x = (x | x << 2) & 0x9249249;
    return x;
}

struct PosPtrs
{
    const double *x, *y, *z;
};

template <int D>
__device__ __forceinline__ void load_pos(const PosPtrs &p, uint32_t i, double *x)
{
    x[0] = p.x[i];
    x[1] = p.y[i];
    if (D == 3)
        x[2] = p.z[i];
}

template <int D>
__global__ void __launch_bounds__(kThreads) k_cell_keys(PosPtrs pos, uint32_t n, MeshParams m, uint32_t *keys,
                                                        uint32_t *vals, uint32_t *counts)
{
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
        return;
    double x[3];
    load_pos<D>(pos, i, x);
    int c[3];
    cell_index<D>(m, x, c);
    uint32_t lin = linear_cell<D>(m, c);
    keys[i] = lin;
    vals[i] = i;
    atomicAdd(&counts[lin], 1u);
}

template <int D>
__global__ void __launch_bounds__(kThreads) k_morton_keys(PosPtrs pos, uint32_t n, MeshParams m, uint32_t *keys,
                                                          uint32_t *vals)
{
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
        return;
    double x[3];
    load_pos<D>(pos, i, x);
    int c[3];
    cell_index<D>(m, x, c);
    uint32_t key = morton_code((uint32_t)c[0]) | (morton_code((uint32_t)c[1]) << 1);
    if (D == 3)
        key |= morton_code((uint32_t)c[2]) << 2;
    keys[i] = key;
    vals[i] = i;
}

__global__ void k_fill_u32(uint32_t *a, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = v;
}
__global__ void k_iota_u32(uint32_t *a, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = (uint32_t)i;
}

//---------------------------------------------------------------------------
// neighbour relations: gather form of UpdateRelation<P,Inner<>> /
// UpdateRelation<P,Contact<>> (update_body_relation.hpp:36-288) over the
// one-cell search box of NeighborSearch (cell_linked_list.hpp:113-167); cells
// visited l -> m -> n, particles in cell-list order.
//---------------------------------------------------------------------------
struct RelationArgs
{
    PosPtrs src, tar;
    uint32_t n_src;
    MeshParams m;
    const uint32_t *cell_offset, *particle_index;
    double inv_h;
    int inner;
    uint32_t *neighbor_size;
    const uint32_t *particle_offset;
    uint32_t *neighbor_index;
};

template <int D, bool FILL>
__global__ void __launch_bounds__(kThreads) k_relation(RelationArgs a)
{
    uint32_t i = xcd_block_index() * kThreads + threadIdx.x;
    if (i >= a.n_src)
        return;
    double xi[3];
    load_pos<D>(a.src, i, xi);
    int c[3], lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    cell_index<D>(a.m, xi, c);
#pragma unroll
    for (int k = 0; k < D; ++k)
    {
        lo[k] = max(0, c[k] - 1);
        hi[k] = min(a.m.all_cells[k], c[k] + 2);
    }
    uint32_t count = 0;
    uint32_t w = FILL ? a.particle_offset[i] : 0u;
    // with the last axis fastest, the cells of one row of the search box are
    // consecutive, so their particles form one contiguous range
    const int m_lo = D == 3 ? lo[1] : 0, m_hi = D == 3 ? hi[1] : 1;
    for (int l = lo[0]; l < hi[0]; ++l)
        for (int mm = m_lo; mm < m_hi; ++mm)
        {
            uint32_t first, last;
            if (D == 2)
            {
                first = (uint32_t)(l * a.m.all_cells[1] + lo[1]);
                last = (uint32_t)(l * a.m.all_cells[1] + hi[1]);
            }
            else
            {
                first = (uint32_t)(l * a.m.all_cells[1] * a.m.all_cells[2] + mm * a.m.all_cells[2] + lo[2]);
                last = first + (uint32_t)(hi[2] - lo[2]);
            }
            uint32_t s0 = a.cell_offset[first], s1 = a.cell_offset[last];
            for (uint32_t s = s0; s < s1; ++s)
            {
                uint32_t j = a.particle_index[s];
                if (a.inner && j == i)
                    continue;
                double xj[3];
                load_pos<D>(a.tar, j, xj);
                // NeighborCriterion: (inv_h * (x_i - x_j)).squaredNorm() < 4
                double s_0 = a.inv_h * (xi[0] - xj[0]);
                double s_1 = a.inv_h * (xi[1] - xj[1]);
                double r2 = s_0 * s_0 + s_1 * s_1;
                if (D == 3)
                {
                    double s_2 = a.inv_h * (xi[2] - xj[2]);
                    r2 = r2 + s_2 * s_2;
                }
                if (r2 < 4.0)
                {
                    if (FILL)
                        a.neighbor_index[w++] = j;
                    else
                        ++count;
                }
            }
        }
    if (!FILL)
        a.neighbor_size[i] = count;
}

//---------------------------------------------------------------------------
// gathers of the particle sort (UpdateSortableVariables,
// base_configuration_dynamics.h:74-128): out[i] = in[perm[i]]
//---------------------------------------------------------------------------
template <class E>
__global__ void k_gather(const E *in, E *out, const uint32_t *perm, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = in[perm[i]];
}
__global__ void k_sorted_id(const uint32_t *original_id, uint32_t *sorted_id, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        sorted_id[original_id[i]] = i;
}
} // namespace

//=============================================================================
// host launchers
//=============================================================================
sphx_status fill_iota(sphx_context *ctx, uint32_t *a, size_t n)
{
    if (n == 0)
        return SPHX_OK;
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, a, n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

sphx_status exclusive_scan_u32(sphx_context *ctx, const uint32_t *in, uint32_t *out, size_t n,
                               uint32_t *total_host)
{
    if (n == 0)
    {
        if (total_host)
            *total_host = 0;
        return SPHX_OK;
    }
    uint32_t ntiles = grid_for(n, kTile);
    sphx_status s = ctx->scan_blocks.reserve((size_t)ntiles * sizeof(uint32_t));
    if (s)
        return s;
    uint32_t *sums = (uint32_t *)ctx->scan_blocks.ptr;
    unsigned long long *tot = ctx->scalars_d + SLOT_TOTAL;
    hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kThreads), 0, ctx->stream, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, ctx->stream, sums, ntiles, tot);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kThreads), 0, ctx->stream, in, out, n, sums);
    SPHX_HIP(hipGetLastError());
    if (total_host)
    {
        if ((s = read_scalars(ctx, SLOT_TOTAL, 1)))
            return s;
        *total_host = (uint32_t)ctx->scalars_h[SLOT_TOTAL];
    }
    return SPHX_OK;
}

sphx_status radix_sort_pairs(sphx_context *ctx, uint32_t *keys, uint32_t *vals, size_t n, int bits)
{
    if (n <= 1 || bits <= 0)
        return SPHX_OK;
    sphx_status s;
    if ((s = ctx->sort_keys.reserve(n * sizeof(uint32_t))) || (s = ctx->sort_vals.reserve(n * sizeof(uint32_t))))
        return s;
    uint32_t ntiles = grid_for(n, kTile);
    size_t nhist = (size_t)kBuckets * ntiles;
    if ((s = ctx->sort_hist.reserve(nhist * sizeof(uint32_t))))
        return s;
    uint32_t *hist = (uint32_t *)ctx->sort_hist.ptr;
    uint32_t *k0 = keys, *v0 = vals;
    uint32_t *k1 = (uint32_t *)ctx->sort_keys.ptr, *v1 = (uint32_t *)ctx->sort_vals.ptr;
    int passes = (bits + kRadixBits - 1) / kRadixBits;
    for (int p = 0; p < passes; ++p)
    {
        int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kThreads), 0, ctx->stream, k0, n, shift, hist, ntiles);
        if ((s = exclusive_scan_u32(ctx, hist, hist, nhist, nullptr)))
            return s;
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kThreads), 0, ctx->stream, k0, v0, k1, v1, n, shift,
                           hist, ntiles);
        SPHX_HIP(hipGetLastError());
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    if (k0 != keys)
    {
        SPHX_HIP(hipMemcpyAsync(keys, k0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
        SPHX_HIP(hipMemcpyAsync(vals, v0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
    }
    return SPHX_OK;
}

static PosPtrs pos_ptrs(const sphx_body *b) { return PosPtrs{b->pos[0], b->pos[1], b->pos[2]}; }

static int bits_for(uint64_t max_value)
{
    int b = 0;
    while (b < 32 && (max_value >> b) != 0)
        ++b;
    return b;
}

// UpdateCellLinkedList<P,RealBody>::exec (update_cell_linked_list.hpp:75-106):
// per-cell counts + exclusive scan give CellOffset; a stable sort of the cell
// keys gives ParticleIndex in ascending particle order inside each cell.
sphx_status cell_linked_list_update(sphx_cell_linked_list *cll)
{
    sphx_body *b = cll->body;
    sphx_context *ctx = b->ctx;
    const size_t nc = cll->ncells;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nc + 1, 256)), dim3(256), 0, ctx->stream, cll->cell_offset,
                       nc + 1, 0u);
    if (b->n > 0)
    {
        // counts accumulate in cell_offset, then scan in place
        if (b->D == 2)
            hipLaunchKernelGGL(k_cell_keys<2>, dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                               pos_ptrs(b), b->n, cll->mesh, cll->keys, cll->particle_index, cll->cell_offset);
        else
            hipLaunchKernelGGL(k_cell_keys<3>, dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                               pos_ptrs(b), b->n, cll->mesh, cll->keys, cll->particle_index, cll->cell_offset);
        SPHX_HIP(hipGetLastError());
    }
    sphx_status s = exclusive_scan_u32(ctx, cll->cell_offset, cll->cell_offset, nc + 1, nullptr);
    if (s)
        return s;
    if ((s = radix_sort_pairs(ctx, cll->keys, cll->particle_index, b->n, cll->key_bits)))
        return s;
    cll->built = true;
    return SPHX_OK;
}

static RelationArgs relation_args(sphx_relation *rel)
{
    RelationArgs a;
    a.src = pos_ptrs(rel->source);
    a.tar = pos_ptrs(rel->target);
    a.n_src = rel->source->n;
    a.m = rel->cll->mesh;
    a.cell_offset = rel->cll->cell_offset;
    a.particle_index = rel->cll->particle_index;
    a.inv_h = rel->inv_h;
    a.inner = rel->inner ? 1 : 0;
    a.neighbor_size = rel->neighbor_size;
    a.particle_offset = rel->particle_offset;
    a.neighbor_index = rel->neighbor_index;
    return a;
}

sphx_status relation_count(sphx_relation *rel)
{
    sphx_context *ctx = rel->source->ctx;
    uint32_t n = rel->source->n;
    if (n == 0)
        return SPHX_OK;
    RelationArgs a = relation_args(rel);
    if (rel->source->D == 2)
        hipLaunchKernelGGL((k_relation<2, false>), dim3(grid_for(n, kThreads)), dim3(kThreads), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL((k_relation<3, false>), dim3(grid_for(n, kThreads)), dim3(kThreads), 0, ctx->stream, a);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

sphx_status relation_fill(sphx_relation *rel)
{
    sphx_context *ctx = rel->source->ctx;
    uint32_t n = rel->source->n;
    if (n == 0)
        return SPHX_OK;
    RelationArgs a = relation_args(rel);
    if (rel->source->D == 2)
        hipLaunchKernelGGL((k_relation<2, true>), dim3(grid_for(n, kThreads)), dim3(kThreads), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL((k_relation<3, true>), dim3(grid_for(n, kThreads)), dim3(kThreads), 0, ctx->stream, a);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// ParticleSortCK::prepareSequence + sort (particle_sort_ck.hpp:61-90): on
// return b->sequence holds the sorted Morton keys and b->perm the permutation
sphx_status morton_sequence(sphx_body *b, sphx_cell_linked_list *cll)
{
    sphx_context *ctx = b->ctx;
    if (b->n == 0)
        return SPHX_OK;
    if (b->D == 2)
        hipLaunchKernelGGL(k_morton_keys<2>, dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                           pos_ptrs(b), b->n, cll->mesh, b->sequence, b->perm);
    else
        hipLaunchKernelGGL(k_morton_keys<3>, dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                           pos_ptrs(b), b->n, cll->mesh, b->sequence, b->perm);
    SPHX_HIP(hipGetLastError());
    int max_axis_cells = 0;
    for (int k = 0; k < b->D; ++k)
        max_axis_cells = std::max(max_axis_cells, cll->mesh.all_cells[k]);
    int axis_bits = std::min(10, bits_for((uint64_t)std::max(0, max_axis_cells - 1)));
    // MortonCode places bit k of axis a at bit 3k + a in 2D as in 3D
    return radix_sort_pairs(ctx, b->sequence, b->perm, b->n, 3 * axis_bits);
}

template <class E>
static sphx_status gather_into(sphx_context *ctx, E *field, const uint32_t *perm, uint32_t n)
{
    sphx_status s = ctx->gather_tmp.reserve((size_t)n * sizeof(E));
    if (s)
        return s;
    E *tmp = (E *)ctx->gather_tmp.ptr;
    hipLaunchKernelGGL(k_gather<E>, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, field, tmp, perm, n);
    SPHX_HIP(hipGetLastError());
    SPHX_HIP(hipMemcpyAsync(field, tmp, (size_t)n * sizeof(E), hipMemcpyDeviceToDevice, ctx->stream));
    return SPHX_OK;
}

sphx_status gather_u32(sphx_context *ctx, uint32_t *field, const uint32_t *perm, size_t n)
{
    return n ? gather_into(ctx, field, perm, (uint32_t)n) : SPHX_OK;
}
sphx_status gather_real(sphx_body *b, void **field, const uint32_t *perm)
{
    if (!b->n)
        return SPHX_OK;
    if (b->P == P_FP64)
        return gather_into(b->ctx, (double *)*field, perm, b->n);
    return gather_into(b->ctx, (float *)*field, perm, b->n);
}
sphx_status gather_double(sphx_body *b, double **field, const uint32_t *perm)
{
    return b->n ? gather_into(b->ctx, *field, perm, b->n) : SPHX_OK;
}
sphx_status gather_rec4(sphx_body *b, void **field, const uint32_t *perm)
{
    if (!b->n)
        return SPHX_OK;
    if (b->P == P_FP64)
        return gather_into(b->ctx, (Rec4<double> *)*field, perm, b->n);
    return gather_into(b->ctx, (Rec4<float> *)*field, perm, b->n);
}
sphx_status sorted_id_update(sphx_body *b)
{
    if (!b->n)
        return SPHX_OK;
    hipLaunchKernelGGL(k_sorted_id, dim3(grid_for(b->n, 256)), dim3(256), 0, b->ctx->stream, b->original_id,
                       b->sorted_id, b->n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}
} // namespace sphx
