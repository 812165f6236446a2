// config_kernels.hip -- configuration dynamics on gfx950: exclusive scan, LDS
// radix sort, cell linked list, neighbour relations, particle sort gathers.
//
// Everything here is integer/index work whose results must be bit-exact with
// the CPU path on identical positions, so floating-point contraction is off
// for this translation unit: cell indices and the neighbour criterion are
// evaluated with the reference's separate IEEE operations.
#pragma clang fp contract(off)

#include "sphx_internal.hpp"

#include <cstring>
#include <mutex>

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
__device__ __forceinline__ void scan_reduce_body(const uint32_t *in, size_t n, uint32_t *sums, uint32_t blk)
{
    size_t base = (size_t)blk * kTile + (size_t)threadIdx.x * kItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n)
            s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0)
        sums[blk] = tot;
}
__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t *in, size_t n, uint32_t *sums)
{
    scan_reduce_body(in, n, sums, blockIdx.x);
}

// single workgroup: exclusive scan of the tile sums in place; grand total to *total
__device__ __forceinline__ void scan_sums_body(uint32_t *sums, uint32_t nsums, unsigned long long *total)
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
__global__ void __launch_bounds__(kThreads) k_scan_sums(uint32_t *sums, uint32_t nsums,
                                                        unsigned long long *total)
{
    scan_sums_body(sums, nsums, total);
}

__device__ __forceinline__ void scan_apply_body(const uint32_t *in, uint32_t *out, size_t n, const uint32_t *sums,
                                                uint32_t blk)
{
    size_t base = (size_t)blk * kTile + (size_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = sums[blk] + block_exclusive_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n)
        {
            out[base + k] = run;
            run += v[k];
        }
}
__global__ void __launch_bounds__(kThreads) k_scan_apply(const uint32_t *in, uint32_t *out, size_t n,
                                                         const uint32_t *sums)
{
    scan_apply_body(in, out, n, sums, blockIdx.x);
}

// the whole scan of at most one tile in one launch; the total to *total
__global__ void __launch_bounds__(kThreads) k_scan_one_tile(const uint32_t *in, uint32_t *out, size_t n,
                                                            unsigned long long *total)
{
    const size_t base = (size_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, &tot); // its barriers order the reads before the writes
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (base + k < n)
        {
            out[base + k] = run;
            run += v[k];
        }
    if (threadIdx.x == 0)
        *total = tot;
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
// cell keys and sort keys (the cell index of a position: cell_index, linear_cell)
//---------------------------------------------------------------------------
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


template <int D>
__device__ __forceinline__ void cell_keys_body(const PosPtrs &pos, uint32_t n, const MeshParams &m, uint32_t *keys,
                                               uint32_t *counts, uint32_t blk)
{
    const uint32_t i = blk * kThreads + threadIdx.x, lane = threadIdx.x & 63u;
    uint32_t lin = 0xffffffffu;
    if (i < n)
    {
        double x[3];
        load_pos<D>(pos, i, x);
        int c[3];
        cell_index<D>(m, x, c);
        lin = linear_cell<D>(m, c);
        keys[i] = lin;
    }
    // one atomic per run of equal cells in the wave (particles are in cell
    // order but for those that moved since the last sort)
    const uint32_t prev = (uint32_t)__shfl_up((int)lin, 1u, 64);
    const bool head = lane == 0u || prev != lin;
    const unsigned long long heads = __ballot(head);
    const unsigned long long above = lane == 63u ? 0ull : heads & (~0ull << (lane + 1u));
    const uint32_t next = above ? (uint32_t)__ffsll((long long)above) - 1u : 64u;
    if (head && i < n)
        atomicAdd(&counts[lin], next - lane);
}
template <int D>
__global__ void __launch_bounds__(kThreads) k_cell_keys(PosPtrs pos, uint32_t n, MeshParams m, uint32_t *keys,
                                                        uint32_t *counts)
{
    cell_keys_body<D>(pos, n, m, keys, counts, blockIdx.x);
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
    // float pre-filter of the criterion: squared distance of the float copies
    // below thr_lo is a neighbour, at or above thr_hi is not, in between the
    // double criterion decides; valid for |x_i| <= pre_xmax on every axis
    const Rec4<float> *src_f, *tar_f;
    float thr_lo, thr_hi;
    double pre_xmax;
    uint32_t n_src;
    MeshParams m;
    const uint32_t *cell_offset, *particle_index;
    double inv_h;
    int inner;
    uint32_t *neighbor_size;
    const uint32_t *chunk_offset;
    uint32_t *slice;
    // per target cell: 1 if a target particle lies within the cell's search
    // box (null: not computed); sources in other cells have no neighbours
    const uint8_t *near;
};

// near[c] of a cell linked list: any particle in the +-1 cell search box of c
template <int D>
__global__ void k_near_flags(const uint32_t *cell_offset, MeshParams m, uint8_t *near, uint32_t ncells)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells)
        return;
    int cc[3];
    if (D == 3)
    {
        cc[2] = (int)(c % (uint32_t)m.all_cells[2]);
        cc[1] = (int)((c / (uint32_t)m.all_cells[2]) % (uint32_t)m.all_cells[1]);
        cc[0] = (int)(c / ((uint32_t)m.all_cells[2] * (uint32_t)m.all_cells[1]));
    }
    else
    {
        cc[1] = (int)(c % (uint32_t)m.all_cells[1]);
        cc[0] = (int)(c / (uint32_t)m.all_cells[1]);
    }
    bool any = false;
    for (int dl = -1; dl <= 1; ++dl)
        for (int dm = -1; dm <= (D == 3 ? 1 : -1); ++dm)
        {
            const int l = cc[0] + dl;
            if (l < 0 || l >= m.all_cells[0])
                continue;
            if (D == 3)
            {
                const int mm = cc[1] + dm;
                if (mm < 0 || mm >= m.all_cells[1])
                    continue;
                const int lo2 = max(0, cc[2] - 1), hi2 = min(m.all_cells[2], cc[2] + 2);
                const uint32_t first = (uint32_t)(l * m.all_cells[1] * m.all_cells[2] + mm * m.all_cells[2] + lo2);
                any |= cell_offset[first + (uint32_t)(hi2 - lo2)] > cell_offset[first];
            }
            else
            {
                const int lo1 = max(0, cc[1] - 1), hi1 = min(m.all_cells[1], cc[1] + 2);
                any |= cell_offset[(uint32_t)(l * m.all_cells[1] + hi1)] > cell_offset[(uint32_t)(l * m.all_cells[1] + lo1)];
            }
        }
    near[c] = any ? 1 : 0;
}

// slices -> CSR NeighborIndex (the reference's layout), for reads
__global__ void k_slice_to_csr(const uint32_t *count, const uint32_t *chunk, const uint32_t *slice,
                               const uint32_t *offset, uint32_t *csr, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t base = chunk[i >> 6] + (i & 63u), o = offset[i];
    for (uint32_t k = 0; k < count[i]; ++k)
        csr[o + k] = slice[base + (k << 6)];
}

constexpr uint32_t kEmpty = 0xffffffffu;

//---------------------------------------------------------------------------
// one-pass relation build: a single candidate scan per source particle writes
// the neighbour slices (fixed kmax rows per 64-particle chunk).
// Rows of a wave are assembled in an LDS ring buffer and written out whole, so
// the list writes are coalesced rows instead of scattered 4-byte stores.
//---------------------------------------------------------------------------
constexpr uint32_t kRing = 16;      // ring rows per wave
constexpr uint32_t kFlushEvery = 8; // candidate iterations between flushes


struct BuildArgs
{
    RelationArgs r;
    uint32_t kmax;
    unsigned long long *needed; // max neighbour count seen when it exceeds kmax
};

template <int D>
__global__ void __launch_bounds__(kThreads) k_relation_build(BuildArgs b)
{
    const RelationArgs &a = b.r;
    __shared__ uint32_t ring[kThreads / 64][kRing][64];
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t blk = xcd_block_index();
    const uint32_t i = blk * kThreads + t;
    const bool valid = i < a.n_src;
    double xi[3] = {0.0, 0.0, 0.0};
    int c[3] = {0, 0, 0};
    Rec4<float> xf = {0.0f, 0.0f, 0.0f, 0.0f};
    bool exact_lane = false;
    if (valid)
    {
        load_pos<D>(a.src, i, xi);
        cell_index<D>(a.m, xi, c);
        xf = a.src_f[i];
        for (int k2 = 0; k2 < D; ++k2)
            exact_lane |= !(fabs(xi[k2]) <= a.pre_xmax);
    }
    const size_t row0 = (size_t)(i >> 6) * b.kmax * 64u + lane;
    uint32_t k = 0;       // neighbours found by this lane
    uint32_t flushed = 0; // rows of this wave already written (wave-uniform)
    auto flush = [&](bool last) {
        const uint32_t mine = min(k, b.kmax);
        const uint32_t kmax_w = wave_max_u32(valid ? mine : 0u);
        const uint32_t kmin_w = wave_min_u32(valid ? mine : 0xffffffffu);
        uint32_t target = last ? kmax_w : min(kmin_w, kmax_w);
        // keep room for the next kFlushEvery appends of the fastest lane
        if (kmax_w + kFlushEvery > flushed + kRing)
            target = max(target, kmax_w + kFlushEvery - kRing);
        for (uint32_t r = flushed; r < target; ++r)
            a.slice[row0 + (size_t)r * 64u] = ring[wave][r % kRing][lane];
        flushed = max(flushed, target);
    };
    // one candidate: float pre-filter, exact double criterion in the margin
    // shell, append to the ring
    auto test = [&](uint32_t j, const Rec4<float> &pf) {
        if (a.inner && j == i)
            return;
        const float fx = xf.x - pf.x, fy = xf.y - pf.y, fz = D == 3 ? xf.z - pf.z : 0.0f;
        float r2f = fx * fx + fy * fy;
        if (D == 3)
            r2f = r2f + fz * fz;
        double r2 = r2f < a.thr_lo ? 0.0 : 8.0; // decided by the pre-filter
        if (exact_lane || (r2f >= a.thr_lo && r2f < a.thr_hi))
        {
            double xj[3];
            load_pos<D>(a.tar, j, xj);
            // NeighborCriterion: (inv_h * (x_i - x_j)).squaredNorm() < 4
            const double s_0 = a.inv_h * (xi[0] - xj[0]);
            const double s_1 = a.inv_h * (xi[1] - xj[1]);
            r2 = s_0 * s_0 + s_1 * s_1;
            if (D == 3)
            {
                const double s_2 = a.inv_h * (xi[2] - xj[2]);
                r2 = r2 + s_2 * s_2;
            }
        }
        if (!(r2 < 4.0))
            return;
        if (k < b.kmax)
        {
            if (k < flushed)
                a.slice[row0 + (size_t)k * 64u] = j;
            else
                ring[wave][k % kRing][lane] = j;
        }
        ++k;
    };
    constexpr uint32_t kBatch = 4;
    static_assert(kFlushEvery % kBatch == 0, "flush on batch boundaries");
    // the 3x3 (3D) or 3 (2D) cell rows of the search box; the cells of one row
    // are consecutive in the cell list, so their particles are one range
    for (int dl = -1; dl <= 1; ++dl)
        for (int dm = -1; dm <= (D == 3 ? 1 : -1); ++dm)
        {
            uint32_t s0 = 0, s1 = 0;
            if (valid)
            {
                const int l = c[0] + dl;
                if (D == 3)
                {
                    const int mm = c[1] + dm;
                    if (l >= 0 && l < a.m.all_cells[0] && mm >= 0 && mm < a.m.all_cells[1])
                    {
                        const int lo2 = max(0, c[2] - 1), hi2 = min(a.m.all_cells[2], c[2] + 2);
                        const uint32_t first =
                            (uint32_t)(l * a.m.all_cells[1] * a.m.all_cells[2] + mm * a.m.all_cells[2] + lo2);
                        s0 = a.cell_offset[first];
                        s1 = a.cell_offset[first + (uint32_t)(hi2 - lo2)];
                    }
                }
                else if (l >= 0 && l < a.m.all_cells[0])
                {
                    const int lo1 = max(0, c[1] - 1), hi1 = min(a.m.all_cells[1], c[1] + 2);
                    s0 = a.cell_offset[(uint32_t)(l * a.m.all_cells[1] + lo1)];
                    s1 = a.cell_offset[(uint32_t)(l * a.m.all_cells[1] + hi1)];
                }
            }
            // wave-uniform trip count so that the ring can be flushed in the loop
            const uint32_t trip = wave_max_u32(s1 - s0);
            for (uint32_t it = 0; it < trip; it += kBatch)
            {
                uint32_t jj[kBatch];
                Rec4<float> pf[kBatch];
#pragma unroll
                for (uint32_t q = 0; q < kBatch; ++q)
                    jj[q] = s0 + it + q < s1 ? a.particle_index[s0 + it + q] : kEmpty;
#pragma unroll
                for (uint32_t q = 0; q < kBatch; ++q)
                    if (jj[q] != kEmpty)
                        pf[q] = a.tar_f[jj[q]];
#pragma unroll
                for (uint32_t q = 0; q < kBatch; ++q)
                    if (jj[q] != kEmpty)
                        test(jj[q], pf[q]);
                if ((it + kBatch) % kFlushEvery == 0)
                    flush(false);
            }
            flush(false);
        }
    flush(true);
    if (valid)
    {
        a.neighbor_size[i] = min(k, b.kmax);
        if (k > b.kmax)
            atomicMax(b.needed, (unsigned long long)k);
    }
}

//---------------------------------------------------------------------------
// tiled relation build: one wave per 64-particle chunk, four chunks per
// workgroup.  The wave takes the chunk's source particles one after another;
// its 64 lanes hold 64 consecutive candidates of the source's flat candidate
// list (the cell rows of its search box, l -> m order, each row the
// particles of up to three consecutive cells in cell order), and the
// neighbours among them are recorded as one 64-bit lane mask per batch of 64
// candidates, in LDS per source.  Sources in the same cell share their
// candidates, which stay in registers until the cell changes.  The output
// phase walks every lane's own masks in candidate order, one neighbour per
// lane per slice row, so each slice row is written whole.
//---------------------------------------------------------------------------
// one wave per workgroup: a wave's LDS and slot are released as soon as it
// ends, not with the slowest wave of a larger group (faster than 2, 4 or 8)
constexpr uint32_t kTileWaves = 1;
constexpr uint32_t kMaskSlots = 15;  // candidate batches of 64 per source: at most 960 candidates (7.5 KB of masks per wave)
constexpr float kFarAway = 1.0e30f; // position of an unused candidate slot (squared distance overflows to inf)

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, uint32_t col)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)col);
}
__device__ __forceinline__ int lane_i32(int v, uint32_t col) { return __builtin_amdgcn_readlane(v, (int)col); }
__device__ __forceinline__ float lane_f32(float v, uint32_t col)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)col));
}
__device__ __forceinline__ double lane_f64(double v, uint32_t col)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), (int)col);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), (int)col);
    return __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
}
// lane mask of a per-lane condition (the compare result itself, no int round trip)
__device__ __forceinline__ unsigned long long lane_mask(bool b) { return __builtin_amdgcn_ballot_w64(b); }
__device__ __forceinline__ bool lane_in(unsigned long long m, uint32_t lane) { return (m >> lane) & 1ull; }

// v[idx] for a per-lane idx as a chain of selects kept in registers (the
// empty asm stops the chain from being turned into a private-memory table)
template <int R>
__device__ __forceinline__ uint32_t pick(const uint32_t (&v)[R], uint32_t idx)
{
    uint32_t out = v[0];
#pragma unroll
    for (int r = 1; r < R; ++r)
    {
        out = idx == (uint32_t)r ? v[r] : out;
        __asm__ volatile("" : "+v"(out));
    }
    return out;
}

template <int D, int W>
__device__ __forceinline__ void relation_tile_body(const BuildArgs &b, const uint32_t blk)
{
    constexpr int R = D == 3 ? 9 : 3; // cell rows of the search box, l -> m order
    constexpr int QMAX = 8;           // candidates of a cell kept in registers: 64 * QMAX
    const RelationArgs &a = b.r;
    __shared__ unsigned long long masks[W][64][kMaskSlots];
    __shared__ uint32_t any_dense;
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t chunk = blk * W + wave;
    const uint32_t i = chunk * 64u + lane;
    const bool valid = i < a.n_src;
    const uint32_t kmax = b.kmax;
    unsigned long long(&my_masks)[64][kMaskSlots] = masks[wave];
    if (t == 0)
        any_dense = 0;
    __syncthreads();
    // this lane's source particle: position, cell and its cell-row ranges
    double xi[3] = {0.0, 0.0, 0.0};
    int c[3] = {0, 0, 0};
    Rec4<float> xf = {0.0f, 0.0f, 0.0f, 0.0f};
    bool exact_lane = false;
    uint32_t s0[R], len[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        s0[r] = len[r] = 0;
    if (valid)
    {
        load_pos<D>(a.src, i, xi);
        cell_index<D>(a.m, xi, c);
        xf = a.src_f[i];
        for (int k2 = 0; k2 < D; ++k2)
            exact_lane |= !(fabs(xi[k2]) <= a.pre_xmax);
    }
    // sources whose search box holds no target particle keep empty rows
    if (valid && (!a.near || a.near[linear_cell<D>(a.m, c)]))
    {
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            const int l = c[0] + (D == 3 ? r / 3 : r) - 1;
            if (D == 3)
            {
                const int mm = c[1] + r % 3 - 1;
                if (l >= 0 && l < a.m.all_cells[0] && mm >= 0 && mm < a.m.all_cells[1])
                {
                    const int lo2 = max(0, c[2] - 1), hi2 = min(a.m.all_cells[2], c[2] + 2);
                    const uint32_t first =
                        (uint32_t)(l * a.m.all_cells[1] * a.m.all_cells[2] + mm * a.m.all_cells[2] + lo2);
                    s0[r] = a.cell_offset[first];
                    len[r] = a.cell_offset[first + (uint32_t)(hi2 - lo2)] - s0[r];
                }
            }
            else if (l >= 0 && l < a.m.all_cells[0])
            {
                const int lo1 = max(0, c[1] - 1), hi1 = min(a.m.all_cells[1], c[1] + 2);
                s0[r] = a.cell_offset[(uint32_t)(l * a.m.all_cells[1] + lo1)];
                len[r] = a.cell_offset[(uint32_t)(l * a.m.all_cells[1] + hi1)] - s0[r];
            }
        }
    }
    uint32_t my_total = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
        my_total += len[r];
    // a search box with more candidates than the masks hold goes to the scan kernel
    if (__ballot(my_total > 64u * kMaskSlots) && lane == 0)
        any_dense = 1;
    __syncthreads();
    if (any_dense)
    {
        if (t == 0)
            atomicMax(b.needed, kTooDense);
        return; // the whole block
    }
    const uint32_t first_i = chunk * 64u;
    // the flat candidate list of the current cell: entry f = 64 q + lane
    uint32_t u_s0[R], u_p[R], u_total = 0; // wave-uniform: row starts, running offsets, length
#pragma unroll
    for (int r = 0; r < R; ++r)
        u_s0[r] = u_p[r] = 0;
    // particle index of candidate f of the flat list
    auto locate = [&](uint32_t f) -> uint32_t {
        uint32_t row = 0, off = f;
#pragma unroll
        for (int r = 1; r < R; ++r)
            if (f >= u_p[r])
            {
                row = (uint32_t)r;
                off = f - u_p[r];
            }
        return a.particle_index[pick<R>(u_s0, row) + off];
    };
    uint32_t cj[QMAX];
    float cx[QMAX], cy[QMAX], cz[QMAX];
#pragma unroll
    for (int q = 0; q < QMAX; ++q)
    {
        cj[q] = 0;
        cx[q] = cy[q] = cz[q] = 0.0f;
    }
    uint32_t my_count = 0;
    // sources with candidates (most fluid particles have no wall cell near)
    // sources are taken cell by cell (all of a cell's sources in the chunk
    // share its candidates), so a particle that left its cell's run since the
    // last sort does not make the wave reload the candidates
    unsigned long long todo = __ballot(my_total > 0u);
    unsigned long long group = 0ull; // sources of the current cell not yet done
    while (todo)
    {
        if (!group)
        {
            // the next cell: its sources and its candidates into registers
            const uint32_t lead = (uint32_t)__ffsll((long long)todo) - 1u;
            const int q0 = lane_i32(c[0], lead), q1 = lane_i32(c[1], lead), q2 = D == 3 ? lane_i32(c[2], lead) : 0;
            group = __ballot(c[0] == q0 && c[1] == q1 && (D == 2 || c[2] == q2)) & todo;
            u_total = 0;
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                u_s0[r] = lane_u32(s0[r], lead);
                u_p[r] = u_total;
                u_total += lane_u32(len[r], lead);
            }
            const uint32_t fill = u_total <= 64u * QMAX ? u_total : 0u;
            // slot f = 64 q + lane of the flat list: its row is the last one
            // starting at or before f (an empty row starts where the next
            // does), picked by vector selects over the uniform row starts; the
            // shift from a slot to its particle_index position is mod 2^32
#pragma unroll
            for (int q = 0; q < QMAX; ++q)
            {
                const uint32_t f = 64u * q + lane;
                uint32_t shift = u_s0[0];
#pragma unroll
                for (int r = 1; r < R; ++r)
                    shift = f >= u_p[r] ? u_s0[r] - u_p[r] : shift;
                if (f < fill)
                    cj[q] = a.particle_index[f + shift];
            }
#pragma unroll
            for (int q = 0; q < QMAX; ++q)
            {
                // slots past the list hold a far position: never a neighbour
                Rec4<float> p = {kFarAway, kFarAway, kFarAway, 0.0f};
                if (64u * q + lane < fill)
                    p = a.tar_f[cj[q]];
                cx[q] = p.x;
                cy[q] = p.y;
                cz[q] = p.z;
            }
        }
        const uint32_t col = (uint32_t)__ffsll((long long)group) - 1u;
        group &= group - 1ull;
        todo &= ~(1ull << col);
        const uint32_t si = first_i + col;
        const uint32_t self = a.inner ? si : kEmpty;
        const float sx = lane_f32(xf.x, col), sy = lane_f32(xf.y, col), sz = D == 3 ? lane_f32(xf.z, col) : 0.0f;
        const bool s_exact = lane_u32(exact_lane ? 1u : 0u, col) != 0u;
        double sx0 = 0.0, sx1 = 0.0, sx2 = 0.0; // the source position, read when a shell needs it
        auto source_position = [&]() {
            sx0 = lane_f64(xi[0], col);
            sx1 = lane_f64(xi[1], col);
            sx2 = D == 3 ? lane_f64(xi[2], col) : 0.0;
        };
        // NeighborCriterion: (inv_h * (x_i - x_j)).squaredNorm() < 4, in double
        auto exact = [&](uint32_t j) -> bool {
            double xj[3];
            load_pos<D>(a.tar, j, xj);
            const double s_0 = a.inv_h * (sx0 - xj[0]);
            const double s_1 = a.inv_h * (sx1 - xj[1]);
            double r2 = s_0 * s_0 + s_1 * s_1;
            if (D == 3)
            {
                const double s_2 = a.inv_h * (sx2 - xj[2]);
                r2 = r2 + s_2 * s_2;
            }
            return r2 < 4.0;
        };
        // the pre-filter: below lo_thr a neighbour, below hi_thr the double
        // criterion decides (exact sources: every listed candidate); the margin
        // bounds the rounding of any evaluation order
        const float lo_thr = s_exact ? -1.0f : a.thr_lo;
        const float hi_thr = s_exact ? kFarAway : a.thr_hi;
        if (u_total <= 64u * QMAX)
        {
            // every batch of the cached candidates at once: neighbour and
            // margin-shell lane masks (the scalar unit, shared by the CU's
            // SIMDs, is the busier one here: the source itself is moved out of
            // range by a select instead of a third mask, and the neighbours
            // are counted from the stored masks at the end)
            unsigned long long hit[QMAX], shell[QMAX], any_shell = 0;
#pragma unroll
            for (int q = 0; q < QMAX; ++q)
            {
                const float fx = sx - cx[q], fy = sy - cy[q], fz = D == 3 ? sz - cz[q] : 0.0f;
                float r2f = fx * fx;
                r2f = fmaf(fy, fy, r2f);
                if (D == 3)
                    r2f = fmaf(fz, fz, r2f);
                r2f = cj[q] == self ? kFarAway : r2f; // never below either bound
                hit[q] = lane_mask(r2f < lo_thr);
                shell[q] = lane_mask(r2f < hi_thr) & ~hit[q];
                any_shell |= shell[q];
            }
            if (any_shell)
            {
                source_position();
#pragma unroll
                for (int q = 0; q < QMAX; ++q)
                    if (shell[q])
                    {
                        bool in = false;
                        if (lane_in(shell[q], lane))
                            in = exact(cj[q]);
                        hit[q] |= lane_mask(in); // shell bits are not in hit
                    }
            }
            // lane q stores mask q
            const uint32_t nq = (u_total + 63u) >> 6;
            unsigned long long m = hit[0];
#pragma unroll
            for (int q = 1; q < QMAX; ++q)
                m = lane == (uint32_t)q ? hit[q] : m;
            if (lane < nq)
                my_masks[col][lane] = m;
        }
        else
        {
            // a cell with more candidates than the registers hold: from memory
            source_position();
            for (uint32_t f0 = 0; f0 < u_total; f0 += 64u)
            {
                bool in = false;
                if (f0 + lane < u_total)
                {
                    const uint32_t j = locate(f0 + lane);
                    if (j != self)
                    {
                        const Rec4<float> p = a.tar_f[j];
                        const float fx = sx - p.x, fy = sy - p.y, fz = D == 3 ? sz - p.z : 0.0f;
                        float r2f = fx * fx;
                        r2f = fmaf(fy, fy, r2f);
                        if (D == 3)
                            r2f = fmaf(fz, fz, r2f);
                        in = r2f < lo_thr;
                        if (!in && r2f < hi_thr)
                            in = exact(j);
                    }
                }
                const unsigned long long m = lane_mask(in);
                if (lane == 0)
                    my_masks[col][f0 >> 6] = m;
            }
        }
    }
    // each lane counts its own neighbours from its masks (a source's
    // candidate list is its cell's: my_total entries)
    for (uint32_t q = 0; q < (my_total + 63u) >> 6; ++q)
        my_count += (uint32_t)__popcll(my_masks[lane][q]);
    const uint32_t mine = valid ? min(my_count, kmax) : 0u;
    if (valid)
    {
        a.neighbor_size[i] = mine;
        if (my_count > kmax)
            atomicMax(b.needed, (unsigned long long)my_count);
    }
    // rows out: each lane walks its own masks in candidate order, one
    // neighbour per slice row, so the rows are written whole
    // this lane's flat-list offsets of its cell rows, and for each row the
    // shift from a flat slot to its particle_index position (mod 2^32)
    uint32_t P[R], shift[R];
    {
        uint32_t acc = 0;
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            P[r] = acc;
            shift[r] = s0[r] - acc;
            acc += len[r];
        }
    }
    const size_t row0 = (size_t)chunk * kmax * 64u + lane;
    const uint32_t rows = wave_max_u32(mine);
    uint32_t cur_q = 0;
    unsigned long long cur = mine > 0u ? my_masks[lane][0] : 0ull;
    constexpr uint32_t kOutBatch = 8; // index gathers in flight per lane
    for (uint32_t r0 = 0; r0 < rows; r0 += kOutBatch)
    {
        uint32_t jj[kOutBatch];
#pragma unroll
        for (uint32_t u = 0; u < kOutBatch; ++u)
        {
            jj[u] = kEmpty;
            if (r0 + u < mine)
            {
                while (cur == 0ull) // the lane's next batch with a neighbour (mine bounds the walk)
                    cur = my_masks[lane][++cur_q];
                const uint32_t f = (cur_q << 6) + (uint32_t)(__ffsll((long long)cur) - 1);
                cur &= cur - 1ull;
                // the last row starting at or before f (rows are in flat order)
                uint32_t sh = shift[0];
#pragma unroll
                for (int r = 1; r < R; ++r)
                    sh = f >= P[r] ? shift[r] : sh;
                jj[u] = a.particle_index[f + sh];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kOutBatch; ++u)
            if (r0 + u < rows)
                a.slice[row0 + (size_t)(r0 + u) * 64u] = jj[u];
    }
}

template <int D>
__global__ void __launch_bounds__(kTileWaves * 64) k_relation_tile(BuildArgs b)
{
    relation_tile_body<D, kTileWaves>(b, xcd_block_index());
}

//---------------------------------------------------------------------------
// staging sets of the inner fluid relation, from its finished slices: for each
// block of kStageBlock source particles the distinct neighbours (stage_keys,
// stage_count) and every entry as a 16-bit rank into them (local_slice), or
// kStageOverflow when the set exceeds the capacity.  A hash set in LDS
// collects the distinct indices in segments of 16 consecutive indices (one
// 16-bit occupancy mask per segment), and an entry's rank is its segment's
// base plus the set bits below it.  The passes use a rank only as an LDS
// address, so their results do not depend on it; keeping runs of consecutive
// indices in ascending order makes their staging loads read 256-byte runs and
// lets neighbouring lanes read neighbouring LDS records (fewer bank conflicts)
// without sorting the set.
//---------------------------------------------------------------------------
struct StageArgs
{
    uint32_t n_src, cap;
    const uint32_t *neighbor_size, *chunk_offset, *slice;
    uint32_t *stage_count, *stage_keys;
    uint16_t *local_slice;
};

// segments of the staging set: 16 consecutive particle indices (a group) share
// one segment, found by open addressing on the group id; the segment holds a
// 16-bit occupancy mask, so a group's staged records keep ascending index order
constexpr uint32_t kSegShift = 4;
constexpr uint32_t kNumSeg = 1024;
static_assert(kNumSeg % kThreads == 0, "segments are split evenly over the threads");

__device__ __forceinline__ uint32_t seg_hash(uint32_t group) { return (group * 2654435761u) >> 22; }

__device__ __forceinline__ uint32_t seg_find(const uint32_t *owner, uint32_t group)
{
    uint32_t s = seg_hash(group);
    while (owner[s] != group)
        s = (s + 1u) & (kNumSeg - 1u);
    return s;
}

__device__ __forceinline__ void stage_sets_body(const StageArgs &s, const uint32_t blk)
{
    __shared__ uint32_t owner[kNumSeg];
    __shared__ uint32_t mask[kNumSeg];
    __shared__ uint32_t seg_base[kNumSeg];
    __shared__ uint32_t seg_full;
    const uint32_t t = threadIdx.x;
    const uint32_t i = blk * kThreads + t;
    const bool valid = i < s.n_src;
    for (uint32_t k = t; k < kNumSeg; k += kThreads)
    {
        owner[k] = kEmpty;
        mask[k] = 0;
    }
    if (t == 0)
        seg_full = 0;
    __syncthreads();
    const uint32_t count = valid ? s.neighbor_size[i] : 0u;
    const size_t row0 = valid ? (size_t)s.chunk_offset[i >> 6] + (i & 63u) : 0;
    constexpr uint32_t kInBatch = 16; // row reads in flight per lane
    constexpr uint32_t kLow = (1u << kSegShift) - 1u;
    // a list runs through its candidates in cell order, ascending index within
    // a cell: consecutive entries of one group are merged into one mask and
    // inserted once
    uint32_t g_cur = kEmpty, m_cur = 0;
    auto insert = [&]() {
        uint32_t h = seg_hash(g_cur), probes = 0;
        while (true)
        {
            uint32_t old = owner[h];
            if (old == kEmpty)
                old = atomicCAS(&owner[h], kEmpty, g_cur);
            if (old == kEmpty || old == g_cur)
            {
                atomicOr(&mask[h], m_cur);
                break;
            }
            h = (h + 1u) & (kNumSeg - 1u);
            if (++probes == kNumSeg)
            {
                seg_full = 1;
                break;
            }
        }
    };
    for (uint32_t r0 = 0; r0 < count; r0 += kInBatch)
    {
        uint32_t jj[kInBatch];
#pragma unroll
        for (uint32_t u = 0; u < kInBatch; ++u)
            jj[u] = r0 + u < count ? s.slice[row0 + (size_t)(r0 + u) * 64u] : kEmpty;
#pragma unroll
        for (uint32_t u = 0; u < kInBatch; ++u)
        {
            if (jj[u] == kEmpty)
                continue;
            const uint32_t g = jj[u] >> kSegShift;
            if (g != g_cur)
            {
                if (g_cur != kEmpty)
                    insert();
                g_cur = g;
                m_cur = 0;
            }
            m_cur |= 1u << (jj[u] & kLow);
        }
    }
    if (g_cur != kEmpty)
        insert();
    __syncthreads();
    // compaction in segment order: rank base of each segment, then its keys
    constexpr uint32_t per = kNumSeg / kThreads;
    uint32_t occ = 0;
    for (uint32_t q = 0; q < per; ++q)
        occ += __popc(mask[t * per + q]);
    uint32_t total;
    uint32_t pos = block_exclusive_scan(occ, &total);
    if (total > s.cap || seg_full) // block-uniform: both read after a barrier
    {
        if (t == 0)
            s.stage_count[blk] = kStageOverflow;
        return;
    }
    uint32_t *keys_out = s.stage_keys + (size_t)blk * s.cap;
    for (uint32_t q = 0; q < per; ++q)
    {
        const uint32_t sidx = t * per + q;
        uint32_t m = mask[sidx];
        seg_base[sidx] = pos;
        const uint32_t first = owner[sidx] << kSegShift;
        while (m)
        {
            keys_out[pos++] = first + (uint32_t)(__ffs(m) - 1);
            m &= m - 1u;
        }
    }
    __syncthreads();
    g_cur = kEmpty;
    uint32_t base_cur = 0;
    for (uint32_t r0 = 0; r0 < count; r0 += kInBatch)
    {
        uint32_t jj[kInBatch];
#pragma unroll
        for (uint32_t u = 0; u < kInBatch; ++u)
            jj[u] = r0 + u < count ? s.slice[row0 + (size_t)(r0 + u) * 64u] : kEmpty;
#pragma unroll
        for (uint32_t u = 0; u < kInBatch; ++u)
            if (jj[u] != kEmpty)
            {
                const uint32_t g = jj[u] >> kSegShift;
                if (g != g_cur)
                {
                    const uint32_t sidx = seg_find(owner, g);
                    base_cur = seg_base[sidx];
                    m_cur = mask[sidx];
                    g_cur = g;
                }
                const uint32_t below = m_cur & ((1u << (jj[u] & kLow)) - 1u);
                s.local_slice[row0 + (size_t)(r0 + u) * 64u] = (uint16_t)(base_cur + __popc(below));
            }
    }
    if (t == 0)
        s.stage_count[blk] = total;
}
__global__ void __launch_bounds__(kThreads) k_stage_sets(StageArgs s)
{
    stage_sets_body(s, xcd_block_index());
}

//---------------------------------------------------------------------------
// Batched configuration update over the scenes of a batch (map entries:
// scene or relation << 16 | block of it).  The cell list is a counting sort:
// per-cell counts, a scan, an atomic scatter and then each cell's indices
// put in ascending order, which is the order of the run-based update above
// and of the reference's canonical cell list.
//---------------------------------------------------------------------------
struct CllScene
{
    PosPtrs pos;
    uint32_t n;
    uint32_t ncells;
    MeshParams m;
    uint32_t *keys, *cell_offset, *particle_index;
    uint32_t *cursor;    // ncells
    uint32_t *tile_sums; // scan tiles of ncells + 1
    uint32_t ntiles;
};

__device__ __forceinline__ uint32_t bmap_owner(uint32_t m) { return m >> 16; }
__device__ __forceinline__ uint32_t bmap_block(uint32_t m) { return m & 0xffffu; }

// counts and cursors of the cells to zero
__device__ __forceinline__ void cll_clear_body(const CllScene &a, uint32_t c)
{
    if (c <= a.ncells)
        a.cell_offset[c] = 0u;
    if (c < a.ncells)
        a.cursor[c] = 0u;
}
// particle i at its cell's offset plus a slot taken atomically: one atomic
// per run of equal cells in the wave (particles are in cell order but for
// those that moved since the last sort), the run's lanes in consecutive slots
// from the run's base (every lane of the wave takes part)
__device__ __forceinline__ void cll_scatter_body(const CllScene &a, uint32_t i)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t key = i < a.n ? a.keys[i] : 0xffffffffu;
    const uint32_t prev = (uint32_t)__shfl_up((int)key, 1u, 64);
    const bool head = lane == 0u || prev != key;
    const unsigned long long heads = __ballot(head);
    const unsigned long long above = lane == 63u ? 0ull : heads & (~0ull << (lane + 1u));
    const uint32_t next = above ? (uint32_t)__ffsll((long long)above) - 1u : 64u;
    uint32_t base = 0u;
    if (head && i < a.n)
        base = atomicAdd(&a.cursor[key], next - lane);
    // the run's head: the last head at or below this lane (lane 0 is one)
    const unsigned long long upto = lane == 63u ? ~0ull : (2ull << lane) - 1ull;
    const uint32_t lead = 63u - (uint32_t)__clzll((long long)(heads & upto));
    base = (uint32_t)__shfl((int)base, (int)lead, 64);
    if (i < a.n)
        a.particle_index[a.cell_offset[key] + base + (lane - lead)] = i;
}
// the indices of cell c in ascending order (cells hold a few particles)
// ascending order of N values in registers (bitonic network, N a power of 2)
template <int N>
__device__ __forceinline__ void sort_network(uint32_t (&v)[N])
{
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; ++i)
            {
                const int l = i ^ j;
                if (l > i)
                {
                    const uint32_t x = v[i], y = v[l];
                    const bool up = (i & k) == 0;
                    v[i] = up ? min(x, y) : max(x, y);
                    v[l] = up ? max(x, y) : min(x, y);
                }
            }
}

constexpr uint32_t kCellSortRegs = 32; // cells up to this size are ordered in registers

__device__ __forceinline__ void cll_order_body(const CllScene &a, uint32_t c)
{
    if (c >= a.ncells)
        return;
    const uint32_t lo = a.cell_offset[c], hi = a.cell_offset[c + 1];
    uint32_t *idx = a.particle_index;
    if (hi - lo <= kCellSortRegs)
    {
        // the cell in registers (past its end the largest value): check, and
        // sort only a cell whose particles arrived in several runs
        const uint32_t n = hi - lo;
        uint32_t v[kCellSortRegs];
#pragma unroll
        for (uint32_t k = 0; k < kCellSortRegs; ++k)
            v[k] = k < n ? idx[lo + k] : 0xffffffffu;
        bool sorted = true;
#pragma unroll
        for (uint32_t k = 1; k < kCellSortRegs; ++k)
            sorted = sorted && (k >= n || v[k - 1] < v[k]);
        if (sorted)
            return;
        sort_network(v);
#pragma unroll
        for (uint32_t k = 0; k < kCellSortRegs; ++k)
            if (k < n)
                idx[lo + k] = v[k];
        return;
    }
    // most cells arrive in order (one run of the scatter): check that with
    // independent loads, eight at a time, before sorting
    bool sorted = true;
    uint32_t last = 0u;
    for (uint32_t q0 = lo; q0 < hi && sorted; q0 += 8u)
    {
        uint32_t v[8];
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k)
            v[k] = q0 + k < hi ? idx[q0 + k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k)
            if (q0 + k < hi)
            {
                sorted = sorted && (q0 + k == lo || last < v[k]);
                last = v[k];
            }
    }
    if (sorted)
        return;
    for (uint32_t q = lo + 1; q < hi; ++q)
    {
        const uint32_t v = idx[q];
        uint32_t r = q;
        while (r > lo && idx[r - 1] > v)
        {
            idx[r] = idx[r - 1];
            --r;
        }
        idx[r] = v;
    }
}
__global__ void __launch_bounds__(kThreads) k_cll_clear(CllScene a) { cll_clear_body(a, blockIdx.x * kThreads + threadIdx.x); }
__global__ void __launch_bounds__(kThreads) k_cll_scatter(CllScene a)
{
    cll_scatter_body(a, blockIdx.x * kThreads + threadIdx.x);
}
__global__ void __launch_bounds__(kThreads) k_cll_order(CllScene a) { cll_order_body(a, blockIdx.x * kThreads + threadIdx.x); }

__global__ void __launch_bounds__(kThreads) kb_cll_clear(const CllScene *__restrict__ tab,
                                                         const uint32_t *__restrict__ map,
                                                         const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[blockIdx.x], sc = bmap_owner(m);
    if (!gate[sc])
        return;
    cll_clear_body(tab[sc], bmap_block(m) * kThreads + threadIdx.x);
}

template <int D>
__global__ void __launch_bounds__(kThreads) kb_cll_count(const CllScene *__restrict__ tab,
                                                         const uint32_t *__restrict__ map,
                                                         const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[blockIdx.x], sc = bmap_owner(m);
    if (!gate[sc])
        return;
    const CllScene &a = tab[sc];
    cell_keys_body<D>(a.pos, a.n, a.m, a.keys, a.cell_offset, bmap_block(m));
}

__global__ void __launch_bounds__(kThreads) kb_cll_scan_reduce(const CllScene *__restrict__ tab,
                                                               const uint32_t *__restrict__ map,
                                                               const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[blockIdx.x], sc = bmap_owner(m);
    if (!gate[sc])
        return;
    const CllScene &a = tab[sc];
    scan_reduce_body(a.cell_offset, (size_t)a.ncells + 1, a.tile_sums, bmap_block(m));
}

// one workgroup per scene
__global__ void __launch_bounds__(kThreads) kb_cll_scan_sums(const CllScene *__restrict__ tab,
                                                             const uint32_t *__restrict__ gate)
{
    if (!gate[blockIdx.x])
        return;
    const CllScene &a = tab[blockIdx.x];
    scan_sums_body(a.tile_sums, a.ntiles, nullptr);
}

__global__ void __launch_bounds__(kThreads) kb_cll_scan_apply(const CllScene *__restrict__ tab,
                                                              const uint32_t *__restrict__ map,
                                                              const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[blockIdx.x], sc = bmap_owner(m);
    if (!gate[sc])
        return;
    const CllScene &a = tab[sc];
    scan_apply_body(a.cell_offset, a.cell_offset, (size_t)a.ncells + 1, a.tile_sums, bmap_block(m));
}

__global__ void __launch_bounds__(kThreads) kb_cll_scatter(const CllScene *__restrict__ tab,
                                                           const uint32_t *__restrict__ map,
                                                           const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[blockIdx.x], sc = bmap_owner(m);
    if (!gate[sc])
        return;
    cll_scatter_body(tab[sc], bmap_block(m) * kThreads + threadIdx.x);
}

__global__ void __launch_bounds__(kThreads) kb_cll_order(const CllScene *__restrict__ tab,
                                                         const uint32_t *__restrict__ map,
                                                         const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[blockIdx.x], sc = bmap_owner(m);
    if (!gate[sc])
        return;
    cll_order_body(tab[sc], bmap_block(m) * kThreads + threadIdx.x);
}

__global__ void kb_zero_u64(unsigned long long *p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}

template <int D, int W>
__global__ void __launch_bounds__(W * 64) kb_relation_tile(const BuildArgs *__restrict__ tab,
                                                           const uint32_t *__restrict__ map,
                                                           const uint32_t *__restrict__ rel_scene,
                                                           const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[xcd_block_index()], r = bmap_owner(m);
    if (!gate[rel_scene[r]])
        return;
    relation_tile_body<D, W>(tab[r], bmap_block(m));
}

__global__ void __launch_bounds__(kThreads) kb_stage_sets(const StageArgs *__restrict__ tab,
                                                          const uint32_t *__restrict__ map,
                                                          const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[xcd_block_index()], sc = bmap_owner(m);
    if (!gate[sc] || tab[sc].n_src == 0u)
        return;
    stage_sets_body(tab[sc], bmap_block(m));
}

__global__ void k_chunk_stride(uint32_t *chunk, uint32_t nchunks, uint32_t stride)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= nchunks)
        chunk[c] = c * stride;
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
                               uint32_t *total_host, uint32_t *sums)
{
    if (n == 0)
    {
        if (total_host)
            *total_host = 0;
        return SPHX_OK;
    }
    uint32_t ntiles = grid_for(n, kTile);
    sphx_status s;
    if (!sums)
    {
        if ((s = ctx->scan_blocks.reserve((size_t)ntiles * sizeof(uint32_t))))
            return s;
        sums = (uint32_t *)ctx->scan_blocks.ptr;
    }
    unsigned long long *tot = ctx->scalars_d + SLOT_TOTAL;
    if (ntiles == 1)
        hipLaunchKernelGGL(k_scan_one_tile, dim3(1), dim3(kThreads), 0, ctx->stream, in, out, n, tot);
    else
    {
        hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kThreads), 0, ctx->stream, in, n, sums);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, ctx->stream, sums, ntiles, tot);
        hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kThreads), 0, ctx->stream, in, out, n, sums);
    }
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
    uint32_t ntiles = grid_for(n, kTile);
    size_t nhist = (size_t)kBuckets * ntiles;
    if ((s = ctx->sort_keys.reserve(n * sizeof(uint32_t))) || (s = ctx->sort_vals.reserve(n * sizeof(uint32_t))) ||
        (s = ctx->sort_hist.reserve(nhist * sizeof(uint32_t))))
        return s;
    uint32_t *hist = (uint32_t *)ctx->sort_hist.ptr;
    uint32_t *k1 = (uint32_t *)ctx->sort_keys.ptr, *v1 = (uint32_t *)ctx->sort_vals.ptr;
    uint32_t *k0 = keys, *v0 = vals;
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

// UpdateCellLinkedList<P,RealBody>::exec (update_cell_linked_list.hpp:75-106)
// as a counting sort: per-cell counts (one atomic per run of equal cells in a
// wave) with the cell keys, an exclusive scan giving CellOffset, an atomic
// scatter of the particles to their cells, then each cell's indices put in
// ascending order, so that ParticleIndex is the canonical list whatever order
// the atomics took.  No size is read back to the host.
sphx_status cell_linked_list_update(sphx_cell_linked_list *cll, const CllWork *work)
{
    sphx_body *b = cll->body;
    sphx_context *ctx = b->ctx;
    // the allocation-free form records no profiling events (their creation allocates)
    ProfileScope prof(ctx, work ? nullptr : "cell_linked_list");
    const size_t nc = cll->ncells;
    const uint32_t n = b->n;
    CllScene a{};
    a.pos = pos_ptrs(b);
    a.n = n;
    a.ncells = (uint32_t)nc;
    a.m = cll->mesh;
    a.keys = cll->keys;
    a.cell_offset = cll->cell_offset;
    a.particle_index = cll->particle_index;
    uint32_t *sums = nullptr;
    sphx_status s;
    if (work)
    {
        a.cursor = work->cursor;
        sums = work->scan_sums;
    }
    else
    {
        if ((s = ctx->cll_cursor.reserve(std::max<size_t>(nc, 1) * sizeof(uint32_t))))
            return s;
        a.cursor = (uint32_t *)ctx->cll_cursor.ptr;
    }
    const dim3 blk(kThreads);
    hipLaunchKernelGGL(k_cll_clear, dim3(grid_for(nc + 1, kThreads)), blk, 0, ctx->stream, a);
    if (n > 0)
    {
        // counts accumulate in cell_offset, then scan in place
        if (b->D == 2)
            hipLaunchKernelGGL(k_cell_keys<2>, dim3(grid_for(n, kThreads)), blk, 0, ctx->stream, a.pos, n, cll->mesh,
                               cll->keys, cll->cell_offset);
        else
            hipLaunchKernelGGL(k_cell_keys<3>, dim3(grid_for(n, kThreads)), blk, 0, ctx->stream, a.pos, n, cll->mesh,
                               cll->keys, cll->cell_offset);
    }
    SPHX_HIP(hipGetLastError());
    if ((s = exclusive_scan_u32(ctx, cll->cell_offset, cll->cell_offset, nc + 1, nullptr, sums)))
        return s;
    if (n > 0)
    {
        hipLaunchKernelGGL(k_cll_scatter, dim3(grid_for(n, kThreads)), blk, 0, ctx->stream, a);
        hipLaunchKernelGGL(k_cll_order, dim3(grid_for(nc, kThreads)), blk, 0, ctx->stream, a);
        SPHX_HIP(hipGetLastError());
    }
    cll->built = true;
    ++cll->build_count;
    return SPHX_OK;
}

//---------------------------------------------------------------------------
// Allocation-free cell-list update: the workspace of the update (a cursor
// per cell and the scan's tile sums) lives in one file-scope device array of
// fixed size.  Each device holds its own copy from module load to process end; its
// device address is looked up once per device (noalloc_scratch_init, called
// by the library's context creation with the device current) and only that
// address is handed to kernels and copies.  Calls on one device share the
// array, so calls that overlap on one device race.
//---------------------------------------------------------------------------
namespace
{
constexpr size_t kNoAllocScanSums = (kNoAllocMaxCells + 1 + kTile - 1) / kTile;
constexpr size_t kNoAllocWords = kNoAllocMaxCells + kNoAllocScanSums;
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
uint32_t *g_noalloc_ptr[kMaxDevices] = {};
} // namespace

// a cursor per cell and the scan's tile sums for at most kNoAllocMaxCells cells (16 MiB)
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

sphx_status noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
    {
        set_error("device index beyond the allocation-free workspace table");
        return SPHX_E_ARG;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[device])
        return SPHX_OK;
    void *p = nullptr;
    SPHX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_noalloc_ptr[device] = (uint32_t *)p;
    return SPHX_OK;
}

sphx_status cell_linked_list_update_noalloc(sphx_cell_linked_list *cll)
{
    const int device = cll->body->ctx->device;
    uint32_t *base = nullptr;
    if (device >= 0 && device < kMaxDevices)
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        base = g_noalloc_ptr[device];
    }
    if (!base)
    {
        set_error("allocation-free workspace not set up on this device (context creation does it)");
        return SPHX_E_STATE;
    }
    if (cll->ncells > kNoAllocMaxCells)
    {
        set_error("allocation-free cell-list update: " + std::to_string(cll->ncells) +
                  " cells exceed its workspace (at most " + std::to_string(kNoAllocMaxCells) +
                  " cells); use sphx_update_cell_linked_list");
        return SPHX_E_CAPACITY;
    }
    CllWork w;
    w.cursor = base;                    // one per cell
    w.scan_sums = base + kNoAllocMaxCells; // tile sums of the cell-offset scan
    return cell_linked_list_update(cll, &w);
}

// near flags of a (static) target cell linked list, recomputed after each rebuild
sphx_status cell_near_flags(sphx_cell_linked_list *cll)
{
    sphx_context *ctx = cll->body->ctx;
    if (cll->near && cll->near_build == cll->build_count)
        return SPHX_OK;
    if (!cll->near)
    {
        sphx_status s = dev_alloc(ctx->stream, (void **)&cll->near, (size_t)cll->ncells);
        if (s)
            return s;
    }
    const uint32_t nc = (uint32_t)cll->ncells;
    if (cll->body->D == 2)
        hipLaunchKernelGGL(k_near_flags<2>, dim3(grid_for(nc, 256)), dim3(256), 0, ctx->stream, cll->cell_offset,
                           cll->mesh, cll->near, nc);
    else
        hipLaunchKernelGGL(k_near_flags<3>, dim3(grid_for(nc, 256)), dim3(256), 0, ctx->stream, cll->cell_offset,
                           cll->mesh, cll->near, nc);
    SPHX_HIP(hipGetLastError());
    cll->near_build = cll->build_count;
    return SPHX_OK;
}

// pre-filter bounds: float rounding of positions bounded by |x| <= xmax
// (target mesh extent + 2.1 h) gives a relative error of the squared
// distance below 2.2e-7 xmax/h + 5.6e-7 near the cut-off; the margin is 4x that
void relation_prefilter(const sphx_relation *rel, float *thr_lo, float *thr_hi, double *pre_xmax)
{
    const MeshParams &m = rel->cll->mesh;
    double X = 0.0;
    for (int k = 0; k < rel->source->D; ++k)
        X = std::max(X, std::max(std::fabs(m.lower[k]), std::fabs(m.lower[k] + m.all_cells[k] * m.gs)));
    const double h = 1.0 / rel->inv_h;
    const double xmax = X + 2.1 * h;
    const double margin = 4.0 * (2.2e-7 * xmax / h + 5.6e-7) + 1e-6;
    const double c2 = 4.0 / (rel->inv_h * rel->inv_h);
    *thr_lo = (float)(c2 * (1.0 - margin));
    *thr_hi = (float)(c2 * (1.0 + margin));
    *pre_xmax = X;
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
    a.chunk_offset = rel->chunk_offset;
    a.slice = rel->slice;
    a.near = rel->inner ? nullptr : rel->cll->near;
    a.src_f = rel->source->pos_f32;
    a.tar_f = rel->target->pos_f32;
    relation_prefilter(rel, &a.thr_lo, &a.thr_hi, &a.pre_xmax);
    return a;
}

sphx_status relation_to_csr(sphx_relation *rel)
{
    uint32_t n = rel->source->n;
    if (n == 0)
        return SPHX_OK;
    hipLaunchKernelGGL(k_slice_to_csr, dim3(grid_for(n, 256)), dim3(256), 0, rel->source->ctx->stream,
                       rel->neighbor_size, rel->chunk_offset, rel->slice, rel->particle_offset, rel->neighbor_index, n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

template <int D>
static sphx_status launch_tile(sphx_context *ctx, const BuildArgs &b, uint32_t nchunks)
{
    hipLaunchKernelGGL((k_relation_tile<D>), dim3(grid_for(nchunks, kTileWaves)), dim3(kTileWaves * 64), 0, ctx->stream,
                       b);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// UpdateRelation in one pass: fixed kmax rows per chunk; *needed > kmax on
// return asks the caller to grow kmax and run again, *needed == kTooDense to
// run again with scan = true (a search box with more candidates than the
// tiled kernel's masks hold).
sphx_status relation_build(sphx_relation *rel, uint32_t kmax, unsigned long long *needed, bool scan)
{
    uint32_t n = rel->source->n;
    sphx_context *ctx = rel->source->ctx;
    if (n == 0)
        return SPHX_OK;
    const uint32_t nchunks = (n + 63u) / 64u;
    if (rel->chunk_kmax != kmax) // the chunk offsets depend on the slice width only
    {
        hipLaunchKernelGGL(k_chunk_stride, dim3(grid_for(nchunks + 1, 256)), dim3(256), 0, ctx->stream,
                           rel->chunk_offset, nchunks, kmax * 64u);
        rel->chunk_kmax = kmax;
    }
    BuildArgs b;
    b.r = relation_args(rel);
    b.kmax = kmax;
    b.needed = needed;
    if (!scan)
    {
        ProfileScope prof(ctx, rel->inner ? "relation_build" : "relation_build_contact");
        return rel->source->D == 2 ? launch_tile<2>(ctx, b, nchunks) : launch_tile<3>(ctx, b, nchunks);
    }
    ProfileScope prof(ctx, "relation_scan");
    dim3 g(grid_for(n, kThreads)), blk(kThreads);
    if (rel->source->D == 2)
        hipLaunchKernelGGL((k_relation_build<2>), g, blk, 0, ctx->stream, b);
    else
        hipLaunchKernelGGL((k_relation_build<3>), g, blk, 0, ctx->stream, b);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// staging sets of a staged relation from its current slices
sphx_status relation_stage(sphx_relation *rel)
{
    uint32_t n = rel->source->n;
    sphx_context *ctx = rel->source->ctx;
    if (n == 0 || !rel->staged)
        return SPHX_OK;
    StageArgs a;
    a.n_src = n;
    a.cap = rel->stage_cap;
    a.neighbor_size = rel->neighbor_size;
    a.chunk_offset = rel->chunk_offset;
    a.slice = rel->slice;
    a.stage_count = rel->stage_count;
    a.stage_keys = rel->stage_keys;
    a.local_slice = rel->local_slice;
    ProfileScope prof(ctx, "relation_stage");
    hipLaunchKernelGGL(k_stage_sets, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, ctx->stream, a);
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

//=============================================================================
// batched configuration update (host side)
//=============================================================================
namespace sphx
{
size_t batch_cll_args_bytes() { return sizeof(CllScene); }
uint32_t batch_cll_tiles(const sphx_cell_linked_list *cll) { return grid_for(cll->ncells + 1, kTile); }

void batch_cll_args(void *dst, sphx_cell_linked_list *cll, uint32_t *cursor, uint32_t *tile_sums)
{
    CllScene a{};
    a.pos = pos_ptrs(cll->body);
    a.n = cll->body->n;
    a.ncells = (uint32_t)cll->ncells;
    a.m = cll->mesh;
    a.keys = cll->keys;
    a.cell_offset = cll->cell_offset;
    a.particle_index = cll->particle_index;
    a.cursor = cursor;
    a.tile_sums = tile_sums;
    a.ntiles = batch_cll_tiles(cll);
    std::memcpy(dst, &a, sizeof(a));
}

size_t batch_build_args_bytes() { return sizeof(BuildArgs); }

void batch_build_args(void *dst, sphx_relation *rel, unsigned long long *needed)
{
    BuildArgs b;
    b.r = relation_args(rel);
    b.kmax = rel->kmax;
    b.needed = needed;
    std::memcpy(dst, &b, sizeof(b));
}

size_t batch_stage_args_bytes() { return sizeof(StageArgs); }

void batch_stage_args(void *dst, sphx_relation *rel)
{
    StageArgs a{};
    if (rel && rel->staged && rel->source->n > 0)
    {
        a.n_src = rel->source->n;
        a.cap = rel->stage_cap;
        a.neighbor_size = rel->neighbor_size;
        a.chunk_offset = rel->chunk_offset;
        a.slice = rel->slice;
        a.stage_count = rel->stage_count;
        a.stage_keys = rel->stage_keys;
        a.local_slice = rel->local_slice;
    }
    std::memcpy(dst, &a, sizeof(a));
}

sphx_status batch_cell_linked_list(const BatchConfig &b)
{
    const CllScene *tab = (const CllScene *)b.cll;
    const dim3 blk(kThreads);
    hipLaunchKernelGGL(kb_cll_clear, dim3(b.blocks_cells), blk, 0, b.stream, tab, b.map_cells, b.gate);
    if (b.D == 2)
        hipLaunchKernelGGL(kb_cll_count<2>, dim3(b.blocks_particles), blk, 0, b.stream, tab, b.map_particles, b.gate);
    else
        hipLaunchKernelGGL(kb_cll_count<3>, dim3(b.blocks_particles), blk, 0, b.stream, tab, b.map_particles, b.gate);
    hipLaunchKernelGGL(kb_cll_scan_reduce, dim3(b.blocks_tiles), blk, 0, b.stream, tab, b.map_tiles, b.gate);
    hipLaunchKernelGGL(kb_cll_scan_sums, dim3(b.n_scenes), blk, 0, b.stream, tab, b.gate);
    hipLaunchKernelGGL(kb_cll_scan_apply, dim3(b.blocks_tiles), blk, 0, b.stream, tab, b.map_tiles, b.gate);
    hipLaunchKernelGGL(kb_cll_scatter, dim3(b.blocks_particles), blk, 0, b.stream, tab, b.map_particles, b.gate);
    hipLaunchKernelGGL(kb_cll_order, dim3(b.blocks_cells), blk, 0, b.stream, tab, b.map_cells, b.gate);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

template <int D, int W>
static void launch_batch_tile(const BatchConfig &b)
{
    hipLaunchKernelGGL((kb_relation_tile<D, W>), dim3(b.blocks_chunks), dim3(W * 64), 0, b.stream,
                       (const BuildArgs *)b.build, b.map_chunks, b.rel_scene, b.gate);
}

sphx_status batch_relations(const BatchConfig &b)
{
    hipLaunchKernelGGL(kb_zero_u64, dim3(grid_for(b.n_rel, 64)), dim3(64), 0, b.stream, b.needed, b.n_rel);
    const int w = b.tile_waves;
    if (w != 1 && w != 2 && w != 4)
    {
        set_error("tiled relation build: 1, 2 or 4 waves per workgroup");
        return SPHX_E_ARG;
    }
    if (b.D == 2)
        w == 1 ? launch_batch_tile<2, 1>(b) : w == 2 ? launch_batch_tile<2, 2>(b) : launch_batch_tile<2, 4>(b);
    else
        w == 1 ? launch_batch_tile<3, 1>(b) : w == 2 ? launch_batch_tile<3, 2>(b) : launch_batch_tile<3, 4>(b);
    if (b.stage)
        hipLaunchKernelGGL(kb_stage_sets, dim3(b.blocks_particles), dim3(kThreads), 0, b.stream,
                           (const StageArgs *)b.stage, b.map_particles, b.gate);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}
} // namespace sphx
