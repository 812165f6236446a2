// fluid_kernels.hip -- the weakly-compressible fluid step on gfx950.
//
// One thread per fluid particle; each thread walks its own neighbour list
// (inner, then each wall contact) in the 64-wide slice layout (entry k of the
// 64 particles of a chunk is one contiguous row), gathers the neighbour's
// (x, y, z, Vol) record and pressure or velocity -- from LDS where the
// relation staged the block's distinct neighbours, from global memory
// otherwise -- evaluates the tabulated Wendland C2 gradient on the fly (cubic
// per table interval, table in LDS) and fuses the reference's initialize /
// interact / update stages of one CK dynamics where the data dependencies
// allow it.  The gather stencil has no dense contraction, so MFMA is not
// used; the bound is HBM/L2 traffic and VALU.
//
// Reference: shared_ck/particle_dynamics/fluid_dynamics/acoustic_step_1st_half.hpp:66-180,
// acoustic_step_2nd_half.hpp:33-137, riemann_solver/riemann_solver_ck.hpp:17-69,
// density_regularization.hpp:40-118, fluid_time_step_ck.{h,hpp,cpp}.
#include "sphx_internal.hpp"

#include <cmath>
#include <cstring>
#include <type_traits>

namespace sphx
{
namespace
{
constexpr int kThreads = 256;
constexpr int kMaxWalls = 4;
static_assert(kThreads == (int)kStageBlock, "staging blocks are pass workgroups");
inline unsigned grid_for(size_t n, size_t per) { return (unsigned)((n + per - 1) / per); }

// 1 / |r| of a pair; a coincident pair (r2 = 0) gets a finite value, so that
// its unit vector r * inv_sqrt(r2) is zero as Eigen's normalized() of a zero
// vector, and its distance r2 * inv_sqrt(r2) is zero
template <class T>
__device__ __forceinline__ T inv_sqrt(T r2);
template <>
__device__ __forceinline__ float inv_sqrt<float>(float r2)
{
    return rsqrtf(fmaxf(r2, 1e-30f));
}
template <>
__device__ __forceinline__ double inv_sqrt<double>(double r2)
{
    return 1.0 / sqrt(fmax(r2, 1e-300));
}

__device__ __forceinline__ float frac_part(float x) { return __builtin_amdgcn_fractf(x); }
__device__ __forceinline__ double frac_part(double x) { return __builtin_amdgcn_fract(x); }

template <class T>
struct WallArgs
{
    const Rec4<T> *posvol; // wall (x, y, z|Vol, Vol|0)
    const T *vel[3];       // wall "Velocity" (average velocity of the solid)
    const T *acc[3];       // wall "Acceleration"
    const T *normal[3];    // wall "NormalDirection"
    const T *Vol_ref;      // wall "VolumetricMeasureRef"
    // neighbour lists in 64-wide slices: entry k of particle i at
    // slice[chunk[i / 64] + 64 k + i % 64], count[i] entries
    const uint32_t *chunk, *count, *slice;
    const KernelPoly<T> *poly;
};

template <class T>
struct FluidArgs
{
    uint32_t n;
    const Rec4<T> *posvol;
    Rec4<T> *vel;
    T *force[3];
    T *force_prior[3];
    const T *mass;
    T *rho, *p, *C, *dC, *S;
    const T *Vol_ref;
    double *dpos[3];
    T rho0, c0, p0, h;
    // single-fluid impedance model (riemann_solver_ck.hpp:59-69)
    T rho0c0, inv_rho0c0_sum, inv_rho0c0_ave, rho0c0_geo_ave, inv_c0_ave;
    int riemann;
    const uint32_t *chunk, *count, *slice; // inner neighbour slices (see WallArgs)
    // LDS staging of the inner neighbours (null stage_count: not staged)
    const uint32_t *stage_count, *stage_keys;
    const uint16_t *local_slice;
    uint32_t stage_cap;
    const KernelPoly<T> *poly;
    int n_wall;
    WallArgs<T> wall[kMaxWalls];
    // LinearCorrectionMatrix components (a, b) at corr[a * D + b]; grad_sum =
    // sum_j B_j e_ij dW_ij V_j of the inner neighbours
    T *corr[9];
    T *grad_sum[3];
    // free-surface indication state
    int32_t *indicator, *prev_indicator;
    T *pos_div;
    // viscous force, its previous value in ForcePrior, kernel gradient integral
    T *visc[3], *prev_visc[3], *kgi[3];
    // device-side sub-step control (null: dt is the kernel argument): the
    // acoustic kernels return at once when STEP_ACTIVE is 0 and take STEP_DT
    const unsigned long long *step;
    // max lanes of the time-step reductions (batched launches: each scene's own)
    unsigned long long *lanes;
};

// dt of an acoustic kernel: the argument, or the device-side sub-step's; false
// when the device-side loop has ended and the launch is to do nothing
template <class T>
__device__ __forceinline__ bool step_dt(const FluidArgs<T> &a, double &dt_d)
{
    if (!a.step)
        return true;
    if (a.step[STEP_ACTIVE] == 0ull)
        return false;
    dt_d = __longlong_as_double((long long)a.step[STEP_DT]);
    return true;
}

// per-particle signal speed max(c0 + |v|, sqrt(4 h |F + F_prior| / m)) of
// AcousticTimeStepCK for the max reduction, or +inf when it or the particle's
// compression is not finite: a diverged particle then makes the acoustic time
// step zero (reported as SPHX_E_NONFINITE) instead of dropping out of the max
template <class T>
__device__ __forceinline__ double signal_or_inf(T c_plus_v, T acc_scale, T compression)
{
    if (!isfinite(c_plus_v) || !isfinite(acc_scale) || !isfinite(compression))
        return __builtin_huge_val();
    return (double)dmax(c_plus_v, acc_scale);
}

// position/volume record of particle j
template <class T, int D>
struct PV
{
    T x, y, z, vol;
};
template <class T, int D>
__device__ __forceinline__ PV<T, D> load_pv(const Rec4<T> *a, uint32_t j)
{
    Rec4<T> r = a[j];
    if (D == 3)
        return PV<T, D>{r.x, r.y, r.z, r.w};
    return PV<T, D>{r.x, r.y, T(0), r.z};
}

// compression helpers: FP64 stores C, MIXED stores C - 1
template <int P>
__device__ __forceinline__ typename Prec<P>::T full_C(typename Prec<P>::T stored)
{
    using T = typename Prec<P>::T;
    return Prec<P>::c_minus_one ? stored + T(1) : stored;
}

// a b + c d as fma(a, b, c d), written out: left to the compiler, which of the
// two products it fuses depends on the code around the sum, so the same pair
// evaluated in two neighbour walks (staged records, global records) could
// differ in the last bit; a sum of three is dot3(a, b, c, d, e, f) =
// fma(e, f, fma(a, b, c d))
__device__ __forceinline__ float mad2(float a, float b, float c, float d) { return fmaf(a, b, c * d); }
__device__ __forceinline__ double mad2(double a, double b, double c, double d) { return fma(a, b, c * d); }
__device__ __forceinline__ float dot3(float a, float b, float c, float d, float e, float f)
{
    return fmaf(e, f, mad2(a, b, c, d));
}
__device__ __forceinline__ double dot3(double a, double b, double c, double d, double e, double f)
{
    return fma(e, f, mad2(a, b, c, d));
}

// tabulated kernel polynomial in LDS
template <class T>
struct alignas(4 * sizeof(T)) PolyLDS
{
    T c[kPolyIntervals][4]; // one interval's coefficients are one aligned 16/32-byte LDS read
};

template <class T>
__device__ __forceinline__ void stage_poly(PolyLDS<T> &lds, const KernelPoly<T> *g, bool grad)
{
    for (int k = threadIdx.x; k < kPolyIntervals * 4; k += blockDim.x)
        lds.c[k >> 2][k & 3] = grad ? g->dw[k >> 2][k & 3] : g->w[k >> 2][k & 3];
}

// kernel value or gradient at distance r (x_per_r = inv_h / dq of the relation);
// neighbours have q < 2, so loc stays within the table
template <class T>
__device__ __forceinline__ T eval_poly(const PolyLDS<T> &lds, T r, T x_per_r)
{
    const T x = r * x_per_r;
    const int loc = min((int)x, kPolyIntervals - 1);
    const T f = frac_part(x);
    const T *c = lds.c[loc];
    return ((c[3] * f + c[2]) * f + c[1]) * f + c[0];
}

// walk the neighbour entries of one particle (slice layout: entry k at
// base + 64 k); the indices of a batch are loaded before any of them is used
// so that their latency overlaps
template <class I, class F>
__device__ __forceinline__ void for_each_neighbor(const I *entries, uint32_t base, uint32_t count, F &&f)
{
    constexpr uint32_t kBatch = 8;
    uint32_t k = 0;
    for (; k + kBatch <= count; k += kBatch)
    {
        // one pointer per batch: the 8 loads take immediate offsets (64 entries apart)
        const I *row = entries + base + ((size_t)k << 6);
        uint32_t idx[kBatch];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            idx[q] = row[q << 6];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            f(idx[q]);
    }
    for (; k < count; ++k)
        f((uint32_t)entries[base + (k << 6)]);
}

// Geometry of pair (i, j) handed to the pair functions of the pipelined walk:
// r_ij = x_i - x_j, 1/|r_ij| and the tabulated kernel (value or gradient) at |r_ij|
template <class T>
struct PairGeo
{
    T dx, dy, dz, ir, w;
};

// one interval's cubic of the tabulated kernel, read as one vector from LDS
template <class T>
struct alignas(4 * sizeof(T)) Coef4
{
    T c0, c1, c2, c3;
};

// r_ij = x_i - x_j, 1/|r_ij| and the kernel table's interval and fraction at
// |r_ij|; the kernel is the interval's cubic at the fraction (pair_kernel),
// evaluated once the interval's coefficients are read
template <int D, class T>
__device__ __forceinline__ void pair_geometry(const PV<T, D> &pi, const PV<T, D> &pj, T x_per_r, PairGeo<T> &g,
                                              int &loc, T &f)
{
    g.dx = pi.x - pj.x;
    g.dy = pi.y - pj.y;
    g.dz = D == 3 ? pi.z - pj.z : T(0);
    const T r2 = D == 3 ? dot3(g.dx, g.dx, g.dy, g.dy, g.dz, g.dz) : mad2(g.dx, g.dx, g.dy, g.dy);
    g.ir = inv_sqrt(r2);
    const T x = (r2 * g.ir) * x_per_r;
    loc = min((int)x, kPolyIntervals - 1);
    f = frac_part(x);
}
template <class T>
__device__ __forceinline__ T pair_kernel(const Coef4<T> &c, T f)
{
    return ((c.c3 * f + c.c2) * f + c.c1) * f + c.c0;
}

// The neighbour walk of for_each_neighbor with the latencies overlapped: the
// entries are taken in batches of 8; the list entries of the next batch are
// loaded while the current one is evaluated, and for each batch the records
// are fetched, the distances computed and the 8 kernel intervals requested
// from LDS before any pair is finished.  fetch(entry) returns a record with
// the neighbour's position and volume in .pv (entries past the count fetch
// entry 0, which every non-empty list has, and are not finished); late(entry,
// record) loads the rest of the record, issued with the kernel intervals once
// the positions are no longer needed (so the batch's records and geometry do
// not hold registers at the same time); pair(record, geo) finishes a pair.
// Pairs are finished in list order with the arithmetic of eval_poly, so the
// sums are those of the plain walk.
template <int D, uint32_t KB = 8, class T, class I, class Fetch, class Late, class Pair>
__device__ __forceinline__ void walk_pipelined(const I *entries, uint32_t base, uint32_t count,
                                               const PV<T, D> &pi, const PolyLDS<T> &lds, T x_per_r,
                                               Fetch &&fetch, Late &&late, Pair &&pair)
{
    // pairs per batch: 8; 4 measured slower in double (4 waves per SIMD instead
    // of 2); 6 is faster for the FP32-mixed passes: pass 2 spills at 8 under its
    // register cap, pass 1 at 6 fits five waves per SIMD
    constexpr uint32_t kBatch = KB;
    const uint32_t nbatch = (count + kBatch - 1) / kBatch;
    uint32_t next[kBatch];
    auto load_entries = [&](uint32_t b) {
        const I *row = entries + base + ((size_t)(b * kBatch) << 6);
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            next[q] = b * kBatch + q < count ? (uint32_t)row[q << 6] : 0u;
    };
    // one batch; the last one (FULL false) may hold fewer than kBatch entries
    auto run_batch = [&](uint32_t b, auto full) {
        constexpr bool FULL = decltype(full)::value;
        uint32_t idx[kBatch];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            idx[q] = next[q];
        if (b + 1 < nbatch)
            load_entries(b + 1);
        const uint32_t rem = count - b * kBatch;
        decltype(fetch(0u)) rec[kBatch];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            rec[q] = fetch(idx[q]);
        PairGeo<T> g[kBatch];
        int loc[kBatch];
        T f[kBatch];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            pair_geometry<D>(pi, rec[q].pv, x_per_r, g[q], loc[q], f[q]);
        Coef4<T> c[kBatch];
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            c[q] = *reinterpret_cast<const Coef4<T> *>(lds.c[loc[q]]);
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            late(idx[q], rec[q]);
        // keep the batch's reads issued together (the scheduler would otherwise
        // sink each next to its use and wait on it alone)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (uint32_t q = 0; q < kBatch; ++q)
            if (FULL || q < rem)
            {
                g[q].w = pair_kernel(c[q], f[q]);
                pair(rec[q], g[q]);
            }
    };
    const uint32_t nfull = count / kBatch;
    if (nbatch > 0)
        load_entries(0);
    for (uint32_t b = 0; b < nfull; ++b)
        run_batch(b, std::true_type{});
    if (nfull < nbatch)
        run_batch(nfull, std::false_type{});
}

//---------------------------------------------------------------------------
// AcousticStep1stHalf::InitializeKernel (acoustic_step_1st_half.hpp:66-74)
//---------------------------------------------------------------------------
template <int D, int P>
__device__ __forceinline__ void init_1st_body(const FluidArgs<typename Prec<P>::T> &a, uint32_t blk, double dt_d)
{
    using T = typename Prec<P>::T;
    uint32_t i = blk * kThreads + threadIdx.x;
    if (i >= a.n || !step_dt(a, dt_d))
        return;
    const T dt = T(dt_d);
    T c = a.C[i] + T(0.5) * dt * a.dC[i];
    a.C[i] = c;
    T rho = full_C<P>(c) * a.rho0;
    a.rho[i] = rho;
    a.p[i] = Prec<P>::c_minus_one ? a.p0 * c : a.p0 * (rho / a.rho0 - T(1));
    // its displacement dpos += v dt / 2 is added by pass 1, which reads v anyway
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_init_1st(FluidArgs<typename Prec<P>::T> a, double dt_d)
{
    init_1st_body<D, P>(a, blockIdx.x, dt_d);
}

//---------------------------------------------------------------------------
// AcousticStep2ndHalf::InitializeKernel (acoustic_step_2nd_half.hpp:33-38)
//---------------------------------------------------------------------------
template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_init_2nd(FluidArgs<typename Prec<P>::T> a, double dt_d)
{
    using T = typename Prec<P>::T;
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.n || !step_dt(a, dt_d))
        return;
    const T dt = T(dt_d);
    Rec4<T> v = a.vel[i];
    a.dpos[0][i] += (double)(v.x * dt * T(0.5));
    a.dpos[1][i] += (double)(v.y * dt * T(0.5));
    if (D == 3)
        a.dpos[2][i] += (double)(v.z * dt * T(0.5));
}

// B v for the correction matrix of particle i (components corr[a * D + b])
template <int D, class T>
struct Mat3
{
    T m[D][D];
    __device__ __forceinline__ void load(T *const *c, uint32_t i)
    {
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int q = 0; q < D; ++q)
                m[r][q] = c[r * D + q][i];
    }
    __device__ __forceinline__ void apply(T x, T y, T z, T &ox, T &oy, T &oz) const
    {
        if (D == 3)
        {
            ox = dot3(m[0][0], x, m[0][1], y, m[0][2], z);
            oy = dot3(m[1][0], x, m[1][1], y, m[1][2], z);
            oz = dot3(m[2][0], x, m[2][1], y, m[2][2], z);
        }
        else
        {
            ox = mad2(m[0][0], x, m[0][1], y);
            oy = mad2(m[1][0], x, m[1][1], y);
            oz = T(0);
        }
    }
};

//---------------------------------------------------------------------------
// pass 1: inner + wall interact of AcousticStep1stHalf, then its update
// (v += (F_prior + F) / m dt); optionally the 2nd half's initialize.
// CORR (LinearCorrectionCK): the inner pair term 2 AverageP(B_j p_i, B_i p_j) e_ij
// = k (p_i B_j + p_j B_i) e_ij is summed as p_i grad_sum_i + B_i sum_j p_j dW V_j e_ij
// (grad_sum_i = sum_j B_j e_ij dW V_j is fixed while positions and volumes are),
// the wall term as B_i sum_w (p_i + p_w) dW V_w e.
// The FP32-mixed pass stages at most kPass1Cap records (20 B each: 28 KB and
// the static tables, five blocks a CU) and walks six pairs per batch (94
// VGPRs: five waves per SIMD, no spills); a block with a larger staged set
// reads its neighbours from global memory in this pass.
//---------------------------------------------------------------------------
constexpr uint32_t kPass1Cap = 1400;
static_assert(kPass1Cap <= kStageCap, "pass 1 stages a prefix of the staging capacity");

template <int P>
__host__ __device__ constexpr uint32_t pass1_cap()
{
    return P == P_MIXED ? kPass1Cap : kStageCap;
}

template <int D, int P, bool RIEMANN, bool FUSE_INIT_2ND, bool CORR>
__device__ __forceinline__ void pass1_body(const FluidArgs<typename Prec<P>::T> &a, const uint32_t blk, double dt_d)
{
    using T = typename Prec<P>::T;
    if (!step_dt(a, dt_d)) // block-uniform: before any barrier
        return;
    __shared__ PolyLDS<T> poly_in;
    __shared__ PolyLDS<T> poly_w[kMaxWalls];
    extern __shared__ __align__(32) unsigned char staged_lds[];
    stage_poly(poly_in, a.poly, true);
    for (int k = 0; k < a.n_wall; ++k)
        stage_poly(poly_w[k], a.wall[k].poly, true);
    const uint32_t i = blk * kThreads + threadIdx.x;
    // staged neighbour records: (x, y, z|Vol, Vol|0) and p of the block's distinct neighbours
    Rec4<T> *s_pv = (Rec4<T> *)staged_lds;
    T *s_p = (T *)(s_pv + pass1_cap<P>());
    const uint32_t n_staged = a.stage_count ? a.stage_count[blk] : kStageOverflow;
    const bool staged = n_staged <= pass1_cap<P>(); // kStageOverflow is above any cap
    if (staged)
    {
        const uint32_t *keys = a.stage_keys + (size_t)blk * a.stage_cap;
        for (uint32_t r = threadIdx.x; r < n_staged; r += kThreads)
        {
            const uint32_t j = keys[r];
            s_pv[r] = a.posvol[j];
            s_p[r] = a.p[j];
        }
    }
    __syncthreads();
    if (i >= a.n)
        return;
    const T dt = T(dt_d);
    const T x_per_r = a.poly->x_per_r;
    const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
    const T p_i = a.p[i];
    // inner: force_sum -= 2 AverageP dW V_j e_ij, AverageP = (p_i + p_j)/2 for one fluid
    T fx = 0, fy = 0, fz = 0, cd = 0;
    struct Rec
    {
        PV<T, D> pv;
        T p;
    };
    auto pair = [&](const Rec &rj, const PairGeo<T> &g) {
        const T dWV = g.w * rj.pv.vol;
        const T s = (CORR ? rj.p : p_i + rj.p) * dWV * g.ir;
        fx += s * g.dx;
        fy += s * g.dy;
        if (D == 3)
            fz += s * g.dz;
        if (RIEMANN)
            cd += (p_i - rj.p) * dWV;
    };
    const uint32_t nb = a.chunk[i >> 6] + (i & 63u), nc = a.count[i];
    auto no_late = [](uint32_t, Rec &) {};
    constexpr uint32_t kWalkBatch = P == P_MIXED ? 6u : 8u;
    if (staged)
        walk_pipelined<D, kWalkBatch>(
            a.local_slice, nb, nc, pi, poly_in, x_per_r,
            [&](uint32_t r) { return Rec{load_pv<T, D>(s_pv, r), s_p[r]}; }, no_late, pair);
    else
        walk_pipelined<D, kWalkBatch>(
            a.slice, nb, nc, pi, poly_in, x_per_r,
            [&](uint32_t j) { return Rec{load_pv<T, D>(a.posvol, j), a.p[j]}; }, no_late, pair);
    const T k_avg = T(2) * a.inv_rho0c0_sum * a.rho0c0; // 2 * AverageP / (p_i + p_j)
    const T vol_i = pi.vol;
    T Fx = a.force[0][i], Fy = a.force[1][i], Fz = D == 3 ? a.force[2][i] : T(0);
    Mat3<D, T> Bi;
    if (CORR)
    {
        Bi.load(a.corr, i);
        T bx, by, bz;
        Bi.apply(fx, fy, fz, bx, by, bz);
        fx = p_i * a.grad_sum[0][i] + bx;
        fy = p_i * a.grad_sum[1][i] + by;
        if (D == 3)
            fz = p_i * a.grad_sum[2][i] + bz;
    }
    Fx += -k_avg * fx * vol_i;
    Fy += -k_avg * fy * vol_i;
    if (D == 3)
        Fz += -k_avg * fz * vol_i;
    const T C_i = full_C<P>(a.C[i]);
    T dC = RIEMANN ? cd * a.inv_rho0c0_ave * C_i : T(0);

    // walls: p_w = p_i + rho_i r max(0, (Fp_i/m_i - a_w).(-e)); force -= (p_i+p_w) dW V_w e
    if (a.n_wall > 0)
    {
        const T m_i = a.mass[i], rho_i = a.rho[i];
        const T apx = a.force_prior[0][i] / m_i, apy = a.force_prior[1][i] / m_i;
        const T apz = D == 3 ? a.force_prior[2][i] / m_i : T(0);
        for (int k = 0; k < a.n_wall; ++k)
        {
            const WallArgs<T> &w = a.wall[k];
            const T w_x_per_r = w.poly->x_per_r;
            T wx = 0, wy = 0, wz = 0, wcd = 0;
            const uint32_t mb = w.chunk[i >> 6] + (i & 63u), mc = w.count[i];
            for (uint32_t k2 = 0; k2 < mc; ++k2)
            {
                const uint32_t j = w.slice[mb + (k2 << 6)];
                const PV<T, D> pj = load_pv<T, D>(w.posvol, j);
                const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
                const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
                const T ir = inv_sqrt(r2);
                const T r = r2 * ir;
                const T dWV = eval_poly(poly_w[k], r, w_x_per_r) * pj.vol;
                const T ex = dx * ir, ey = dy * ir, ez = D == 3 ? dz * ir : T(0);
                T face = -((apx - w.acc[0][j]) * ex + (apy - w.acc[1][j]) * ey);
                if (D == 3)
                    face -= (apz - w.acc[2][j]) * ez;
                const T p_w = p_i + rho_i * r * dmax(T(0), face);
                const T s = (p_i + p_w) * dWV;
                wx += s * ex;
                wy += s * ey;
                if (D == 3)
                    wz += s * ez;
                if (RIEMANN)
                    wcd += (p_i - p_w) * dWV;
            }
            if (CORR)
            {
                T bx, by, bz;
                Bi.apply(wx, wy, wz, bx, by, bz);
                wx = bx, wy = by, wz = bz;
            }
            Fx += -wx * vol_i;
            Fy += -wy * vol_i;
            if (D == 3)
                Fz += -wz * vol_i;
            if (RIEMANN)
                dC += wcd * a.inv_rho0c0_ave * C_i;
        }
    }
    // inside the device-side sub-step loop the 2nd half, which sets the force
    // anew, runs next whenever this pass runs and nothing reads the force in
    // between: the store is left out there
    if (!a.step)
    {
        a.force[0][i] = Fx;
        a.force[1][i] = Fy;
        if (D == 3)
            a.force[2][i] = Fz;
    }
    a.dC[i] = dC;
    // UpdateKernel: v += (F_prior + F) / m * dt; before it, the initialize's
    // displacement dpos += v dt / 2 with the velocity it saw (nothing changes
    // v between the two kernels) and, fused, the 2nd half's with the new one:
    // the two additions of the separate kernels, in their order
    const T m_i = a.mass[i];
    Rec4<T> v = a.vel[i];
    double d0 = a.dpos[0][i] + (double)(v.x * dt * T(0.5));
    double d1 = a.dpos[1][i] + (double)(v.y * dt * T(0.5));
    double d2 = D == 3 ? a.dpos[2][i] + (double)(v.z * dt * T(0.5)) : 0.0;
    v.x += (a.force_prior[0][i] + Fx) / m_i * dt;
    v.y += (a.force_prior[1][i] + Fy) / m_i * dt;
    if (D == 3)
        v.z += (a.force_prior[2][i] + Fz) / m_i * dt;
    a.vel[i] = v;
    if (FUSE_INIT_2ND)
    {
        d0 += (double)(v.x * dt * T(0.5));
        d1 += (double)(v.y * dt * T(0.5));
        if (D == 3)
            d2 += (double)(v.z * dt * T(0.5));
    }
    a.dpos[0][i] = d0;
    a.dpos[1][i] = d1;
    if (D == 3)
        a.dpos[2][i] = d2;
}

template <int D, int P, bool RIEMANN, bool FUSE_INIT_2ND, bool CORR>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(P == P_MIXED ? 5 : 1))) k_pass1(FluidArgs<typename Prec<P>::T> a, double dt_d)
{
    pass1_body<D, P, RIEMANN, FUSE_INIT_2ND, CORR>(a, xcd_block_index(), dt_d);
}

// DissipativePJump of the Riemann solvers divided by rho0c0_geo_ave (the
// factor is applied once per particle): u times the limiter of
// AcousticRiemannSolverCK, TruncatedLinear(3) on inv_c0 * max(u, 0), or 1
template <class T>
__device__ __forceinline__ T p_jump_scaled(const FluidArgs<T> &a, T u)
{
    return a.riemann == SPHX_RIEMANN_ACOUSTIC ? u * dmin(T(3) * (a.inv_c0_ave * dmax(u, T(0))), T(1)) : u;
}


//---------------------------------------------------------------------------
// Cell traversal of the inner neighbours, the alternative to stored lists
// (kept for their A/B: sphx_acoustic_step_2nd_half_traversal).  One wave per
// 64-particle chunk.  The cells of the chunk's search boxes (the bounding box
// of its particles' cells grown by one cell each way) are staged in LDS one
// (x, y) column of the box after another (a column's z cells are contiguous
// in cell-list order) as the float record (x, y, z, Vol) and the particle
// index.  Each lane then walks its own search box in the order of the stored
// lists (the nine (x, y) columns x-major, each column's three z cells in
// cell-list order) and takes a candidate by the test of the list build: the
// float pre-filter with its rounding margin, the double criterion in the
// margin shell, the particle itself excluded.  The neighbours and their order
// are those of the lists, so the sums are the same bit for bit.  A chunk whose
// box holds more than kTravCap particles or kTravCols columns reads its
// candidates from global memory.  3D, FP32-mixed (the float record is the
// pre-filter's position).
//---------------------------------------------------------------------------
struct TravArgs
{
    const uint32_t *cell_offset, *particle_index; // the fluid's cell linked list
    MeshParams m;
    PosPtrs pos;          // double positions, for the criterion in the margin shell
    double inv_h, pre_xmax;
    float thr_lo, thr_hi; // squared-distance bounds of the pre-filter
};
constexpr uint32_t kTravCols = 64;  // staged (x, y) columns per chunk at most
constexpr uint32_t kTravCap = 1152; // staged candidates per chunk at most (20 B each)

// NeighborCriterion in double, evaluated as the list build does (no contraction)
__device__ __forceinline__ bool trav_exact(const TravArgs &t, const double *xi, uint32_t j)
{
#pragma clang fp contract(off)
    const double s0 = t.inv_h * (xi[0] - t.pos.x[j]);
    const double s1 = t.inv_h * (xi[1] - t.pos.y[j]);
    const double s2 = t.inv_h * (xi[2] - t.pos.z[j]);
    double r2 = s0 * s0 + s1 * s1;
    r2 = r2 + s2 * s2;
    return r2 < 4.0;
}

// a lane's place in the traversal: its cell and position, the chunk's box
struct TravLane
{
    double xi[3];
    int c[3];
    int lo0, lo1, ny; // box origin (x, y) and its extent in y
    bool staged;      // wave-uniform
};

// stage the chunk's box (every lane of the wave takes part)
__device__ __forceinline__ TravLane trav_stage(const TravArgs &t, const Rec4<float> *posvol, uint32_t i, bool valid,
                                               Rec4<float> *s_pv, uint32_t *s_j, uint32_t *col_src, uint32_t *col_dst)
{
    TravLane L{};
    const uint32_t lane = threadIdx.x & 63u;
    if (valid)
    {
        load_pos<3>(t.pos, i, L.xi);
        cell_index<3>(t.m, L.xi, L.c);
    }
    int lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        lo[k] = max((int)wave_min_u32(valid ? (uint32_t)L.c[k] : ~0u) - 1, 0);
        hi[k] = min((int)wave_max_u32(valid ? (uint32_t)L.c[k] : 0u) + 1, t.m.all_cells[k] - 1);
    }
    L.lo0 = lo[0];
    L.lo1 = lo[1];
    L.ny = hi[1] - lo[1] + 1;
    const uint32_t ncol = (uint32_t)((hi[0] - lo[0] + 1) * L.ny);
    uint32_t src = 0, len = 0;
    if (lane < ncol)
    {
        const int l = lo[0] + (int)lane / L.ny, mm = lo[1] + (int)lane % L.ny;
        const uint32_t first = (uint32_t)((l * t.m.all_cells[1] + mm) * t.m.all_cells[2] + lo[2]);
        src = t.cell_offset[first];
        len = t.cell_offset[first + (uint32_t)(hi[2] - lo[2]) + 1u] - src;
    }
    uint32_t incl = len; // inclusive scan over the lanes
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, off, 64);
        if ((int)lane >= off)
            incl += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
    L.staged = ncol <= kTravCols && total <= kTravCap;
    if (L.staged)
    {
        if (lane < ncol)
        {
            col_src[lane] = src;
            col_dst[lane] = incl - len;
        }
        __syncthreads();
        for (uint32_t f = lane; f < total; f += 64u)
        {
            // the column of slot f: the last one starting at or before f
            uint32_t a = 0, b = ncol;
            while (b - a > 1u)
            {
                const uint32_t mid = (a + b) >> 1;
                if (col_dst[mid] <= f)
                    a = mid;
                else
                    b = mid;
            }
            const uint32_t j = t.particle_index[col_src[a] + (f - col_dst[a])];
            s_pv[f] = posvol[j];
            s_j[f] = j;
        }
    }
    __syncthreads();
    return L;
}

// walk the lane's search box; pair(j, record, geometry) for each neighbour j
template <class Pair>
__device__ __forceinline__ void trav_walk(const TravArgs &t, const TravLane &L, const Rec4<float> *posvol,
                                          uint32_t i, const PV<float, 3> &pi, const PolyLDS<float> &lds,
                                          float x_per_r, const Rec4<float> *s_pv, const uint32_t *s_j,
                                          const uint32_t *col_src, const uint32_t *col_dst, Pair &&pair)
{
    constexpr float kFar = 1.0e30f;
    bool exact_lane = false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        exact_lane |= !(fabs(L.xi[k]) <= t.pre_xmax);
    const float lo_thr = exact_lane ? -1.0f : t.thr_lo;
    const float hi_thr = exact_lane ? kFar : t.thr_hi;
    const int zlo = max(0, L.c[2] - 1), zhi = min(t.m.all_cells[2] - 1, L.c[2] + 1);
    for (int r = 0; r < 9; ++r)
    {
        const int l = L.c[0] + r / 3 - 1, mm = L.c[1] + r % 3 - 1;
        if (l < 0 || l >= t.m.all_cells[0] || mm < 0 || mm >= t.m.all_cells[1])
            continue;
        const uint32_t first = (uint32_t)((l * t.m.all_cells[1] + mm) * t.m.all_cells[2] + zlo);
        const uint32_t s0 = t.cell_offset[first];
        const uint32_t n = t.cell_offset[first + (uint32_t)(zhi - zlo) + 1u] - s0;
        uint32_t base = 0;
        if (L.staged)
        {
            const uint32_t q = (uint32_t)((l - L.lo0) * L.ny + (mm - L.lo1));
            base = col_dst[q] + (s0 - col_src[q]);
        }
        for (uint32_t k = 0; k < n; ++k)
        {
            uint32_t j;
            Rec4<float> p;
            if (L.staged)
            {
                j = s_j[base + k];
                p = s_pv[base + k];
            }
            else
            {
                j = t.particle_index[s0 + k];
                p = posvol[j];
            }
            float fx = pi.x - p.x, fy = pi.y - p.y, fz = pi.z - p.z;
            // the criterion's own copies: its products are not merged with the
            // pair geometry's, which then contracts as in the list walk
            __asm__ volatile("" : "+v"(fx), "+v"(fy), "+v"(fz));
            float r2f = fx * fx;
            r2f = fmaf(fy, fy, r2f);
            r2f = fmaf(fz, fz, r2f);
            bool in = j != i && r2f < lo_thr;
            if (j != i && !in && r2f < hi_thr)
                in = trav_exact(t, L.xi, j);
            if (in)
            {
                const PV<float, 3> pj{p.x, p.y, p.z, p.w};
                PairGeo<float> g;
                int loc;
                float f;
                pair_geometry<3>(pi, pj, x_per_r, g, loc, f);
                g.w = pair_kernel(*reinterpret_cast<const Coef4<float> *>(lds.c[loc]), f);
                pair(j, pj, g);
            }
        }
    }
}

//---------------------------------------------------------------------------
// pass 2: inner + wall interact of AcousticStep2ndHalf, its update
// (C += dt/2 dC, rho = C rho0) and, fused, the per-particle signal speed of
// AcousticTimeStepCK for the next sub-step, max-reduced into the max lanes at max_slot.
//---------------------------------------------------------------------------
// CORR (LinearCorrectionCK): the divergence uses B_i e_ij in place of e_ij
// Pass 2 stages at most kPass2Cap records, as (x, y, z|Vol, Vol|0) plus the
// three velocity components in separate arrays (28 B a record), so that with
// the static tables a block takes at most 40 KB of LDS and four blocks fit a
// CU; a block with a larger staged set reads its neighbours from global
// memory in this pass only.
constexpr uint32_t kPass2Cap = 1400;
static_assert(kPass2Cap <= kStageCap, "pass 2 stages a prefix of the staging capacity");

// FP32-mixed: registers capped for 4 waves per SIMD, which its 40 KB of LDS
// per block allows; FP64 (not staged) keeps its registers
// TRAV: the inner neighbours by the cell traversal (one-wave blocks, 3D,
// FP32-mixed, not staged by the stored lists' sets)
template <int D, int P, bool RIEMANN, bool FUSE_CFL, bool CORR, bool TRAV = false>
__device__ __forceinline__ void pass2_body(const FluidArgs<typename Prec<P>::T> &a, const uint32_t blk, double dt_d,
                                           unsigned long long *max_slot, const TravArgs *trav = nullptr)
{
    using T = typename Prec<P>::T;
    static_assert(!TRAV || (D == 3 && P == P_MIXED && !CORR), "the traversal is 3D FP32-mixed, uncorrected");
    constexpr uint32_t NT = TRAV ? 64u : kThreads; // threads per block
    if (!step_dt(a, dt_d)) // block-uniform: before any barrier
        return;
    __shared__ PolyLDS<T> poly_in;
    __shared__ PolyLDS<T> poly_w[kMaxWalls];
    __shared__ double wave_red[kThreads / 64];
    extern __shared__ __align__(32) unsigned char staged_lds[];
    stage_poly(poly_in, a.poly, true);
    for (int k = 0; k < a.n_wall; ++k)
        stage_poly(poly_w[k], a.wall[k].poly, true);
    const uint32_t i = blk * NT + threadIdx.x;
    // staged neighbour records: (x, y, z|Vol, Vol|0) and the velocity components
    Rec4<T> *s_pv = (Rec4<T> *)staged_lds;
    // (no opaque base as in pass 1: its extra register makes this pass spill,
    // 45% slower measured)
    T *s_vx = (T *)(s_pv + kPass2Cap), *s_vy = s_vx + kPass2Cap, *s_vz = s_vy + kPass2Cap;
    const uint32_t n_staged = a.stage_count ? a.stage_count[blk] : kStageOverflow;
    const bool staged = n_staged <= kPass2Cap; // kStageOverflow is above any cap
    if (staged)
    {
        const uint32_t *keys = a.stage_keys + (size_t)blk * a.stage_cap;
        for (uint32_t r = threadIdx.x; r < n_staged; r += kThreads)
        {
            const uint32_t j = keys[r];
            s_pv[r] = a.posvol[j];
            const Rec4<T> v = a.vel[j];
            s_vx[r] = v.x;
            s_vy[r] = v.y;
            if (D == 3)
                s_vz[r] = v.z;
        }
    }
    // the traversal's staged box: records, particle indices, column starts
    __shared__ uint32_t trav_cols[TRAV ? 2 * kTravCols : 1];
    Rec4<float> *t_pv = (Rec4<float> *)staged_lds;
    uint32_t *t_j = (uint32_t *)(t_pv + kTravCap);
    TravLane tl{};
    if constexpr (TRAV)
        tl = trav_stage(*trav, a.posvol, i, i < a.n, t_pv, t_j, trav_cols, trav_cols + kTravCols);
    __syncthreads();
    double signal = -1.0;
    if (i < a.n)
    {
        const T dt = T(dt_d);
        const T x_per_r = a.poly->x_per_r;
        const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
        const Rec4<T> vi = a.vel[i];
        Mat3<D, T> Bi;
        if (CORR)
            Bi.load(a.corr, i);
        // divergence: 2 (v_i - v_ave).e dW V = (v_i - v_j).e dW V for one fluid
        T div = 0, px = 0, py = 0, pz = 0;
        struct Rec
        {
            PV<T, D> pv;
            Rec4<T> v;
        };
        auto pair = [&](const Rec &rj, const PairGeo<T> &g) {
            const Rec4<T> &vj = rj.v;
            const T dWV = g.w * rj.pv.vol;
            const T ir = g.ir;
            const T ex = g.dx * ir, ey = g.dy * ir, ez = D == 3 ? g.dz * ir : T(0);
            T u = mad2(vi.x - vj.x, ex, vi.y - vj.y, ey);
            if (D == 3)
                u = fma(vi.z - vj.z, ez, u);
            if (CORR)
            {
                T bx, by, bz;
                Bi.apply(ex, ey, ez, bx, by, bz);
                T uc = mad2(vi.x - vj.x, bx, vi.y - vj.y, by);
                if (D == 3)
                    uc = fma(vi.z - vj.z, bz, uc);
                div += uc * dWV;
            }
            else
                div += u * dWV;
            if (RIEMANN)
            {
                const T s = p_jump_scaled(a, u) * dWV;
                px += s * ex;
                py += s * ey;
                if (D == 3)
                    pz += s * ez;
            }
        };
        const uint32_t nb = a.chunk[i >> 6] + (i & 63u), nc = a.count[i];
        // whole records up front: loading the velocities late (as pass 1 loads
        // its pressures) makes this pass spill under its register cap
        auto no_late = [](uint32_t, Rec &) {};
        constexpr uint32_t kWalkBatch = P == P_MIXED ? 6u : 8u;
        if constexpr (TRAV)
            trav_walk(*trav, tl, a.posvol, i, pi, poly_in, x_per_r, t_pv, t_j, trav_cols, trav_cols + kTravCols,
                      [&](uint32_t j, const PV<T, D> &pj, const PairGeo<T> &g) { pair(Rec{pj, a.vel[j]}, g); });
        else if (staged)
            walk_pipelined<D, kWalkBatch>(
                a.local_slice, nb, nc, pi, poly_in, x_per_r,
                [&](uint32_t r) {
                    return Rec{load_pv<T, D>(s_pv, r), Rec4<T>{s_vx[r], s_vy[r], D == 3 ? s_vz[r] : T(0), T(0)}};
                },
                no_late, pair);
        else
            walk_pipelined<D, kWalkBatch>(
                a.slice, nb, nc, pi, poly_in, x_per_r,
                [&](uint32_t j) { return Rec{load_pv<T, D>(a.posvol, j), a.vel[j]}; }, no_late, pair);
        const T C_i = full_C<P>(a.C[i]);
        const T vol_i = pi.vol;
        T dC = a.dC[i] + div * C_i;
        const T zv = a.rho0c0_geo_ave * vol_i;
        T Fx = px * zv, Fy = py * zv, Fz = D == 3 ? pz * zv : T(0);
        for (int k = 0; k < a.n_wall; ++k)
        {
            const WallArgs<T> &w = a.wall[k];
            const T w_x_per_r = w.poly->x_per_r;
            T wdiv = 0, wx = 0, wy = 0, wz = 0;
            const uint32_t mb = w.chunk[i >> 6] + (i & 63u), mc = w.count[i];
            for (uint32_t k2 = 0; k2 < mc; ++k2)
            {
                const uint32_t j = w.slice[mb + (k2 << 6)];
                const PV<T, D> pj = load_pv<T, D>(w.posvol, j);
                const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
                const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
                const T ir = inv_sqrt(r2);
                const T dWV = eval_poly(poly_w[k], r2 * ir, w_x_per_r) * pj.vol;
                const T ex = dx * ir, ey = dy * ir, ez = D == 3 ? dz * ir : T(0);
                const T vdx = T(2) * (vi.x - w.vel[0][j]), vdy = T(2) * (vi.y - w.vel[1][j]);
                const T vdz = D == 3 ? T(2) * (vi.z - w.vel[2][j]) : T(0);
                T bex = ex, bey = ey, bez = ez;
                if (CORR)
                    Bi.apply(ex, ey, ez, bex, bey, bez);
                T ve = vdx * bex + vdy * bey;
                if (D == 3)
                    ve += vdz * bez;
                wdiv += ve * dWV;
                if (RIEMANN)
                {
                    const T nx = w.normal[0][j], ny = w.normal[1][j], nz = D == 3 ? w.normal[2][j] : T(0);
                    T en = ex * nx + ey * ny;
                    if (D == 3)
                        en += ez * nz;
                    const T sg = en < T(0) ? T(-1) : (en > T(0) ? T(1) : T(0));
                    const T fnx = sg * nx, fny = sg * ny, fnz = sg * nz;
                    T u = vdx * fnx + vdy * fny;
                    if (D == 3)
                        u += vdz * fnz;
                    const T s = p_jump_scaled(a, u) * dWV;
                    wx += s * fnx;
                    wy += s * fny;
                    if (D == 3)
                        wz += s * fnz;
                }
            }
            dC += wdiv * C_i;
            Fx += wx * zv;
            Fy += wy * zv;
            if (D == 3)
                Fz += wz * zv;
        }
        a.dC[i] = dC;
        a.force[0][i] = Fx;
        a.force[1][i] = Fy;
        if (D == 3)
            a.force[2][i] = Fz;
        // UpdateKernel: C += 0.5 dt dC; rho = C rho0
        T c = a.C[i] + T(0.5) * dt * dC;
        a.C[i] = c;
        a.rho[i] = full_C<P>(c) * a.rho0;
        if (FUSE_CFL)
        {
            // AcousticTimeStepCK::reduce (fluid_time_step_ck.hpp:51-57)
            const T m_i = a.mass[i];
            T gx = Fx + a.force_prior[0][i], gy = Fy + a.force_prior[1][i];
            T gz = D == 3 ? Fz + a.force_prior[2][i] : T(0);
            T fn = sqrt(D == 3 ? gx * gx + gy * gy + gz * gz : gx * gx + gy * gy);
            T acc_scale = sqrt(T(4) * a.h * fn / m_i);
            T vn = sqrt(D == 3 ? vi.x * vi.x + vi.y * vi.y + vi.z * vi.z : vi.x * vi.x + vi.y * vi.y);
            signal = signal_or_inf(a.c0 + vn, acc_scale, c);
        }
    }
    if (FUSE_CFL)
    {
        double m = wave_max(signal);
        if ((threadIdx.x & 63) == 0)
            wave_red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double b = wave_red[0];
            for (int w = 1; w < (int)(NT / 64); ++w)
                b = fmax(b, wave_red[w]);
            atomic_max_spread(max_slot, b);
        }
    }
}

template <int D, int P, bool RIEMANN, bool FUSE_CFL, bool CORR>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(P == P_MIXED ? 4 : 1)))
k_pass2(FluidArgs<typename Prec<P>::T> a, double dt_d, unsigned long long *max_slot)
{
    pass2_body<D, P, RIEMANN, FUSE_CFL, CORR>(a, xcd_block_index(), dt_d, max_slot);
}

// pass 2 with the inner neighbours by the cell traversal: one 64-particle chunk per block
template <bool RIEMANN>
__global__ void __launch_bounds__(64) k_pass2_traverse(FluidArgs<float> a, TravArgs t, double dt_d)
{
    pass2_body<3, P_MIXED, RIEMANN, false, false, true>(a, xcd_block_index(), dt_d, nullptr, &t);
}

//---------------------------------------------------------------------------
// CompressionSummation<Inner<>> + <Contact<>> (density_regularization.hpp:40-83)
//---------------------------------------------------------------------------
template <int D, int P>
__device__ __forceinline__ void summation_body(const FluidArgs<typename Prec<P>::T> &a, const uint32_t blk)
{
    using T = typename Prec<P>::T;
    __shared__ PolyLDS<T> poly_in;
    __shared__ PolyLDS<T> poly_w[kMaxWalls];
    extern __shared__ __align__(32) unsigned char staged_lds[];
    stage_poly(poly_in, a.poly, false);
    for (int k = 0; k < a.n_wall; ++k)
        stage_poly(poly_w[k], a.wall[k].poly, false);
    const uint32_t i = blk * kThreads + threadIdx.x;
    // staged neighbour records: (x, y, z, Vol_ref) of the block's distinct neighbours
    Rec4<T> *s_pv = (Rec4<T> *)staged_lds;
    const uint32_t n_staged = a.stage_count ? a.stage_count[blk] : kStageOverflow;
    const bool staged = n_staged != kStageOverflow;
    if (staged)
    {
        const uint32_t *keys = a.stage_keys + (size_t)blk * a.stage_cap;
        for (uint32_t r = threadIdx.x; r < n_staged; r += kThreads)
        {
            const uint32_t j = keys[r];
            Rec4<T> rec = a.posvol[j];
            if (D == 3)
                rec.w = a.Vol_ref[j];
            else
                rec.z = a.Vol_ref[j];
            s_pv[r] = rec;
        }
    }
    __syncthreads();
    if (i >= a.n)
        return;
    const T x_per_r = a.poly->x_per_r;
    const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
    T s = a.poly->w0 * a.Vol_ref[i];
    // W_ij Vref_j (pj.vol carries Vref_j)
    auto pair = [&](const PV<T, D> &pj) {
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
        const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
        const T r = sqrt(r2);
        s += eval_poly(poly_in, r, x_per_r) * pj.vol;
    };
    const uint32_t nb = a.chunk[i >> 6] + (i & 63u), nc = a.count[i];
    if (staged)
        for_each_neighbor(a.local_slice, nb, nc, [&](uint32_t r) { pair(load_pv<T, D>(s_pv, r)); });
    else
        for_each_neighbor(a.slice, nb, nc, [&](uint32_t j) {
            PV<T, D> pj = load_pv<T, D>(a.posvol, j);
            pj.vol = a.Vol_ref[j];
            pair(pj);
        });
    for (int k = 0; k < a.n_wall; ++k)
    {
        const WallArgs<T> &w = a.wall[k];
        const T w_x_per_r = w.poly->x_per_r;
        const uint32_t mb = w.chunk[i >> 6] + (i & 63u), mc = w.count[i];
        for (uint32_t k2 = 0; k2 < mc; ++k2)
        {
            const uint32_t j = w.slice[mb + (k2 << 6)];
            const PV<T, D> pj = load_pv<T, D>(w.posvol, j);
            const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
            const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
            const T r = sqrt(r2);
            s += eval_poly(poly_w[k], r, w_x_per_r) * w.Vol_ref[j];
        }
    }
    a.S[i] = s;
}

// FP32-mixed: six blocks a CU (26 KB of staged records each, 66 VGPRs)
template <int D, int P>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(P == P_MIXED ? 6 : 1)))
k_summation(FluidArgs<typename Prec<P>::T> a)
{
    summation_body<D, P>(a, xcd_block_index());
}

//---------------------------------------------------------------------------
// LinearCorrectionMatrix<Inner<WithUpdate>> + <Contact<>> (kernel_correction_ck.hpp:37-91).
// B_raw = -sum_j r_ij (x) grad W_ij V_j = -sum_j (dW_ij V_j / |r_ij|) r_ij r_ij^T is
// symmetric, so six (three in 2D) sums are kept; the update (det, Tikhonov
// inverse, blend with I) runs in double per particle in both precisions.
//---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ double det_d(const double (&m)[3][3])
{
    if (D == 2)
        return m[0][0] * m[1][1] - m[1][0] * m[0][1];
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[1][0] * (m[0][1] * m[2][2] - m[0][2] * m[2][1]) +
           m[2][0] * (m[0][1] * m[1][2] - m[0][2] * m[1][1]);
}
// adjugate / determinant, the determinant from the first-column cofactors
template <int D>
__device__ __forceinline__ void inverse_d(const double (&x)[3][3], double (&r)[3][3])
{
    if (D == 2)
    {
        const double inv_det = 1.0 / (x[0][0] * x[1][1] - x[1][0] * x[0][1]);
        r[0][0] = x[1][1] * inv_det;
        r[1][0] = -x[1][0] * inv_det;
        r[0][1] = -x[0][1] * inv_det;
        r[1][1] = x[0][0] * inv_det;
        return;
    }
    double cof[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            cof[i][j] = x[i1][j1] * x[i2][j2] - x[i1][j2] * x[i2][j1];
        }
    const double det = cof[0][0] * x[0][0] + cof[1][0] * x[1][0] + cof[2][0] * x[2][0];
    const double inv_det = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[i][j] = cof[j][i] * inv_det;
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_correction(FluidArgs<typename Prec<P>::T> a, double alpha)
{
    using T = typename Prec<P>::T;
    __shared__ PolyLDS<T> poly_in;
    __shared__ PolyLDS<T> poly_w[kMaxWalls];
    extern __shared__ __align__(32) unsigned char staged_lds[];
    stage_poly(poly_in, a.poly, true);
    for (int k = 0; k < a.n_wall; ++k)
        stage_poly(poly_w[k], a.wall[k].poly, true);
    const uint32_t blk = xcd_block_index();
    const uint32_t i = blk * kThreads + threadIdx.x;
    Rec4<T> *s_pv = (Rec4<T> *)staged_lds;
    const uint32_t n_staged = a.stage_count ? a.stage_count[blk] : kStageOverflow;
    const bool staged = n_staged != kStageOverflow;
    if (staged)
    {
        const uint32_t *keys = a.stage_keys + (size_t)blk * a.stage_cap;
        for (uint32_t r = threadIdx.x; r < n_staged; r += kThreads)
            s_pv[r] = a.posvol[keys[r]];
    }
    __syncthreads();
    if (i >= a.n)
        return;
    const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
    // b_ab = sum dW V / |r| r_a r_b (B_raw = -b)
    T bxx = 0, bxy = 0, bxz = 0, byy = 0, byz = 0, bzz = 0;
    auto acc = [&](const PV<T, D> &pj, const PolyLDS<T> &poly, T x_per_r) {
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
        const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
        const T ir = inv_sqrt(r2);
        const T s = eval_poly(poly, r2 * ir, x_per_r) * pj.vol * ir;
        const T sx = s * dx, sy = s * dy;
        bxx += sx * dx;
        bxy += sx * dy;
        byy += sy * dy;
        if (D == 3)
        {
            const T sz = s * dz;
            bxz += sx * dz;
            byz += sy * dz;
            bzz += sz * dz;
        }
    };
    const T x_per_r = a.poly->x_per_r;
    const uint32_t nb = a.chunk[i >> 6] + (i & 63u), nc = a.count[i];
    if (staged)
        for_each_neighbor(a.local_slice, nb, nc, [&](uint32_t r) { acc(load_pv<T, D>(s_pv, r), poly_in, x_per_r); });
    else
        for_each_neighbor(a.slice, nb, nc, [&](uint32_t j) { acc(load_pv<T, D>(a.posvol, j), poly_in, x_per_r); });
    for (int k = 0; k < a.n_wall; ++k)
    {
        const WallArgs<T> &w = a.wall[k];
        const uint32_t mb = w.chunk[i >> 6] + (i & 63u), mc = w.count[i];
        for (uint32_t k2 = 0; k2 < mc; ++k2)
            acc(load_pv<T, D>(w.posvol, w.slice[mb + (k2 << 6)]), poly_w[k], w.poly->x_per_r);
    }
    // UpdateKernel::update (kernel_correction_ck.hpp:58-65), in double
    double B[3][3] = {{-(double)bxx, -(double)bxy, -(double)bxz},
                      {-(double)bxy, -(double)byy, -(double)byz},
                      {-(double)bxz, -(double)byz, -(double)bzz}};
    const double det = det_d<D>(B);
    const double det_sqr = fmax(alpha - det, 0.0);
    // inverseTikhonov: (B^T B + 1e-8 I)^-1 B^T (common/vector_functions.h:198-203)
    double A[3][3], Ai[3][3];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int q = 0; q < D; ++q)
        {
            double t = B[0][r] * B[0][q];
#pragma unroll
            for (int k = 1; k < D; ++k)
                t += B[k][r] * B[k][q];
            A[r][q] = t + (r == q ? 1.0e-8 : 0.0);
        }
    inverse_d<D>(A, Ai);
    const double weight = det / (det + det_sqr);
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int q = 0; q < D; ++q)
        {
            double t = Ai[r][0] * B[q][0];
#pragma unroll
            for (int k = 1; k < D; ++k)
                t += Ai[r][k] * B[q][k];
            a.corr[r * D + q][i] = T(weight * t + (1.0 - weight) * (r == q ? 1.0 : 0.0));
        }
}

// grad_sum_i = sum_j B_j e_ij dW_ij V_j over the inner neighbours (the part of
// the corrected pressure force that does not change within an advection step)
template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_correction_grad(FluidArgs<typename Prec<P>::T> a)
{
    using T = typename Prec<P>::T;
    __shared__ PolyLDS<T> poly_in;
    stage_poly(poly_in, a.poly, true);
    __syncthreads();
    const uint32_t i = xcd_block_index() * kThreads + threadIdx.x;
    if (i >= a.n)
        return;
    const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
    const T x_per_r = a.poly->x_per_r;
    T gx = 0, gy = 0, gz = 0;
    for_each_neighbor(a.slice, a.chunk[i >> 6] + (i & 63u), a.count[i], [&](uint32_t j) {
        const PV<T, D> pj = load_pv<T, D>(a.posvol, j);
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
        const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
        const T ir = inv_sqrt(r2);
        const T dWV = eval_poly(poly_in, r2 * ir, x_per_r) * pj.vol;
        Mat3<D, T> Bj;
        Bj.load(a.corr, j);
        T bx, by, bz;
        Bj.apply(dx * ir, dy * ir, D == 3 ? dz * ir : T(0), bx, by, bz);
        gx += bx * dWV;
        gy += by * dWV;
        if (D == 3)
            gz += bz * dWV;
    });
    a.grad_sum[0][i] = gx;
    a.grad_sum[1][i] = gy;
    if (D == 3)
        a.grad_sum[2][i] = gz;
}

//---------------------------------------------------------------------------
// FreeSurfaceIndicationCK<Inner<WithUpdate>> + <Contact<>>
// (surface_indication/surface_indication_ck.hpp:55-157): position divergence
// -sum dW V r, held at twice the threshold for a particle that neither was nor
// neighbours the previous free surface, then the contact part ...
//---------------------------------------------------------------------------
template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_surface_div(FluidArgs<typename Prec<P>::T> a)
{
    using T = typename Prec<P>::T;
    __shared__ PolyLDS<T> poly_in;
    __shared__ PolyLDS<T> poly_w[kMaxWalls];
    stage_poly(poly_in, a.poly, true);
    for (int k = 0; k < a.n_wall; ++k)
        stage_poly(poly_w[k], a.wall[k].poly, true);
    __syncthreads();
    const uint32_t i = xcd_block_index() * kThreads + threadIdx.x;
    if (i >= a.n)
        return;
    const T threshold = T(0.75) * T(D);
    const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
    auto term = [&](const PV<T, D> &pj, const PolyLDS<T> &poly, T x_per_r) {
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
        const T r = sqrt(D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy);
        return eval_poly(poly, r, x_per_r) * pj.vol * r;
    };
    const uint32_t nb = a.chunk[i >> 6] + (i & 63u), nc = a.count[i];
    T pos_div = 0;
    for_each_neighbor(a.slice, nb, nc,
                      [&](uint32_t j) { pos_div -= term(load_pv<T, D>(a.posvol, j), poly_in, a.poly->x_per_r); });
    if (pos_div < threshold && a.prev_indicator[i] != 1)
    {
        bool near_previous_surface = false;
        for (uint32_t k = 0; k < nc; ++k)
            near_previous_surface |= a.prev_indicator[a.slice[nb + (k << 6)]] == 1;
        if (!near_previous_surface)
            pos_div = T(2) * threshold;
    }
    for (int k = 0; k < a.n_wall; ++k)
    {
        const WallArgs<T> &w = a.wall[k];
        const uint32_t mb = w.chunk[i >> 6] + (i & 63u), mc = w.count[i];
        T wall_div = 0;
        for (uint32_t k2 = 0; k2 < mc; ++k2)
            wall_div -= term(load_pv<T, D>(w.posvol, w.slice[mb + (k2 << 6)]), poly_w[k], w.poly->x_per_r);
        pos_div += wall_div;
    }
    a.pos_div[i] = pos_div;
}

// ... and the update: Indicator = 0 when the divergence exceeds the threshold and
// no inner neighbour closer than h is below it, else 1; PreviousSurfaceIndicator = Indicator
template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_surface_update(FluidArgs<typename Prec<P>::T> a)
{
    using T = typename Prec<P>::T;
    const uint32_t i = xcd_block_index() * kThreads + threadIdx.x;
    if (i >= a.n)
        return;
    const T threshold = T(0.75) * T(D);
    int32_t indicator = 1;
    if (a.pos_div[i] > threshold)
    {
        const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
        bool very_near_surface = false;
        for_each_neighbor(a.slice, a.chunk[i >> 6] + (i & 63u), a.count[i], [&](uint32_t j) {
            const PV<T, D> pj = load_pv<T, D>(a.posvol, j);
            const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
            const T r = sqrt(D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy);
            very_near_surface |= (a.pos_div[j] < threshold) & (r < a.h);
        });
        if (!very_near_surface)
            indicator = 0;
    }
    a.indicator[i] = indicator;
    a.prev_indicator[i] = indicator;
}

//---------------------------------------------------------------------------
// ViscousForceCK<Inner<WithUpdate, Viscosity, K>> + <Contact<Wall, Viscosity, K>>
// (viscous_force.hpp:38-104) and the ForcePriorCK update (force_prior_ck.h:55-59):
// F_i = V_i [ sum_j mu_ij (r.(C_i + C_j) e) (v_i - v_j) / (r^2 + 0.01 h^2) dW V_j
//           + sum_w 2 mu (r.C_i e) 2 (v_i - v_w) / (r^2 + 0.01 h^2) dW V_w ],
// C = 1 (CORR false) or the correction matrix; ForcePrior += F - F_previous.
//---------------------------------------------------------------------------
template <int D, int P, bool CORR>
__global__ void __launch_bounds__(kThreads) k_viscous(FluidArgs<typename Prec<P>::T> a, double mu_pair_d,
                                                      double mu_d, double h_sq_d)
{
    using T = typename Prec<P>::T;
    __shared__ PolyLDS<T> poly_in;
    __shared__ PolyLDS<T> poly_w[kMaxWalls];
    stage_poly(poly_in, a.poly, true);
    for (int k = 0; k < a.n_wall; ++k)
        stage_poly(poly_w[k], a.wall[k].poly, true);
    __syncthreads();
    const uint32_t i = xcd_block_index() * kThreads + threadIdx.x;
    if (i >= a.n)
        return;
    const T mu_pair = T(mu_pair_d), mu = T(mu_d), eps = T(0.01) * T(h_sq_d);
    const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
    const Rec4<T> vi = a.vel[i];
    Mat3<D, T> Bi;
    if (CORR)
        Bi.load(a.corr, i);
    T fx = 0, fy = 0, fz = 0;
    const T x_per_r = a.poly->x_per_r;
    for_each_neighbor(a.slice, a.chunk[i >> 6] + (i & 63u), a.count[i], [&](uint32_t j) {
        const PV<T, D> pj = load_pv<T, D>(a.posvol, j);
        const Rec4<T> vj = a.vel[j];
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
        const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
        const T ir = inv_sqrt(r2);
        const T dWV = eval_poly(poly_in, r2 * ir, x_per_r) * pj.vol;
        T s; // r . (C_i + C_j) e
        if (CORR)
        {
            Mat3<D, T> Bj;
            Bj.load(a.corr, j);
            T bx, by, bz, cx, cy, cz;
            const T ex = dx * ir, ey = dy * ir, ez = D == 3 ? dz * ir : T(0);
            Bi.apply(ex, ey, ez, bx, by, bz);
            Bj.apply(ex, ey, ez, cx, cy, cz);
            s = dx * (bx + cx) + dy * (by + cy);
            if (D == 3)
                s += dz * (bz + cz);
        }
        else
            s = T(2) * (r2 * ir);
        const T c = s * mu_pair * dWV / (r2 + eps);
        fx += c * (vi.x - vj.x);
        fy += c * (vi.y - vj.y);
        if (D == 3)
            fz += c * (vi.z - vj.z);
    });
    const T vol_i = pi.vol;
    T Fx = fx * vol_i, Fy = fy * vol_i, Fz = fz * vol_i;
    for (int k = 0; k < a.n_wall; ++k)
    {
        const WallArgs<T> &w = a.wall[k];
        const T w_x_per_r = w.poly->x_per_r;
        T wx = 0, wy = 0, wz = 0;
        const uint32_t mb = w.chunk[i >> 6] + (i & 63u), mc = w.count[i];
        for (uint32_t k2 = 0; k2 < mc; ++k2)
        {
            const uint32_t j = w.slice[mb + (k2 << 6)];
            const PV<T, D> pj = load_pv<T, D>(w.posvol, j);
            const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
            const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
            const T ir = inv_sqrt(r2);
            const T dWV = eval_poly(poly_w[k], r2 * ir, w_x_per_r) * pj.vol;
            T s; // r . C_i e
            if (CORR)
            {
                T bx, by, bz;
                Bi.apply(dx * ir, dy * ir, D == 3 ? dz * ir : T(0), bx, by, bz);
                s = dx * bx + dy * by;
                if (D == 3)
                    s += dz * bz;
            }
            else
                s = r2 * ir;
            const T c = T(2) * s * mu * dWV / (r2 + eps);
            wx += c * (T(2) * (vi.x - w.vel[0][j]));
            wy += c * (T(2) * (vi.y - w.vel[1][j]));
            if (D == 3)
                wz += c * (T(2) * (vi.z - w.vel[2][j]));
        }
        Fx += wx * vol_i;
        Fy += wy * vol_i;
        if (D == 3)
            Fz += wz * vol_i;
    }
    // ForcePriorCK::UpdateKernel::update
    a.visc[0][i] = Fx;
    a.visc[1][i] = Fy;
    a.force_prior[0][i] += Fx - a.prev_visc[0][i];
    a.force_prior[1][i] += Fy - a.prev_visc[1][i];
    a.prev_visc[0][i] = Fx;
    a.prev_visc[1][i] = Fy;
    if (D == 3)
    {
        a.visc[2][i] = Fz;
        a.force_prior[2][i] += Fz - a.prev_visc[2][i];
        a.prev_visc[2][i] = Fz;
    }
}

//---------------------------------------------------------------------------
// KernelGradientIntegral<Inner<K>> + <Contact<Boundary, K>> (kernel_gradient_integral.hpp:37-49,72-84):
// g_i = -sum_j (C_i + C_j) dW V_j e_ij - sum_w 2 C_i dW V_w e_iw
//---------------------------------------------------------------------------
template <int D, int P, bool CORR>
__global__ void __launch_bounds__(kThreads) k_kernel_gradient_integral(FluidArgs<typename Prec<P>::T> a)
{
    using T = typename Prec<P>::T;
    __shared__ PolyLDS<T> poly_in;
    __shared__ PolyLDS<T> poly_w[kMaxWalls];
    stage_poly(poly_in, a.poly, true);
    for (int k = 0; k < a.n_wall; ++k)
        stage_poly(poly_w[k], a.wall[k].poly, true);
    __syncthreads();
    const uint32_t i = xcd_block_index() * kThreads + threadIdx.x;
    if (i >= a.n)
        return;
    const PV<T, D> pi = load_pv<T, D>(a.posvol, i);
    Mat3<D, T> Bi;
    if (CORR)
        Bi.load(a.corr, i);
    T gx = 0, gy = 0, gz = 0;
    const T x_per_r = a.poly->x_per_r;
    for_each_neighbor(a.slice, a.chunk[i >> 6] + (i & 63u), a.count[i], [&](uint32_t j) {
        const PV<T, D> pj = load_pv<T, D>(a.posvol, j);
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
        const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
        const T ir = inv_sqrt(r2);
        const T dWV = eval_poly(poly_in, r2 * ir, x_per_r) * pj.vol;
        const T ex = dx * ir, ey = dy * ir, ez = D == 3 ? dz * ir : T(0);
        if (CORR)
        {
            Mat3<D, T> Bj;
            Bj.load(a.corr, j);
            T bx, by, bz, cx, cy, cz;
            Bi.apply(ex, ey, ez, bx, by, bz);
            Bj.apply(ex, ey, ez, cx, cy, cz);
            gx -= (bx + cx) * dWV;
            gy -= (by + cy) * dWV;
            if (D == 3)
                gz -= (bz + cz) * dWV;
        }
        else
        {
            const T s = T(2) * dWV;
            gx -= s * ex;
            gy -= s * ey;
            if (D == 3)
                gz -= s * ez;
        }
    });
    for (int k = 0; k < a.n_wall; ++k)
    {
        const WallArgs<T> &w = a.wall[k];
        const T w_x_per_r = w.poly->x_per_r;
        T wx = 0, wy = 0, wz = 0;
        const uint32_t mb = w.chunk[i >> 6] + (i & 63u), mc = w.count[i];
        for (uint32_t k2 = 0; k2 < mc; ++k2)
        {
            const uint32_t j = w.slice[mb + (k2 << 6)];
            const PV<T, D> pj = load_pv<T, D>(w.posvol, j);
            const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
            const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
            const T ir = inv_sqrt(r2);
            const T s = T(2) * eval_poly(poly_w[k], r2 * ir, w_x_per_r) * pj.vol;
            T ex = dx * ir, ey = dy * ir, ez = D == 3 ? dz * ir : T(0);
            if (CORR)
                Bi.apply(dx * ir, dy * ir, D == 3 ? dz * ir : T(0), ex, ey, ez);
            wx -= s * ex;
            wy -= s * ey;
            if (D == 3)
                wz -= s * ez;
        }
        gx += wx;
        gy += wy;
        if (D == 3)
            gz += wz;
    }
    a.kgi[0][i] = gx;
    a.kgi[1][i] = gy;
    if (D == 3)
        a.kgi[2][i] = gz;
}

// TransportVelocityCorrectionCK::UpdateKernel::update (transport_velocity_correction_ck.hpp:39-51):
// dpos += coefficient h^2 limiter(h^2 |g|^2) g (h ratio 1), TruncatedLinear(100) or no limiter,
// all particles or those with Indicator 0
template <int D, int P>
__global__ void k_transport_correction(FluidArgs<typename Prec<P>::T> a, double scaling_d, double sq_h_d,
                                       int truncated, int bulk_only)
{
    using T = typename Prec<P>::T;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n || (bulk_only && a.indicator[i] != 0))
        return;
    const T gx = a.kgi[0][i], gy = a.kgi[1][i], gz = D == 3 ? a.kgi[2][i] : T(0);
    const T sq = D == 3 ? gx * gx + gy * gy + gz * gz : gx * gx + gy * gy;
    const T lim = truncated ? dmin(T(100) * (T(sq_h_d) * sq), T(1)) : T(1);
    const T c = T(scaling_d) * lim;
    a.dpos[0][i] += (double)(c * gx);
    a.dpos[1][i] += (double)(c * gy);
    if (D == 3)
        a.dpos[2][i] += (double)(c * gz);
}

//---------------------------------------------------------------------------
// state dynamics
//---------------------------------------------------------------------------
template <int D, int P>
__device__ __forceinline__ void regularize_body(const FluidArgs<typename Prec<P>::T> &a, uint32_t i, int flow)
{
    using T = typename Prec<P>::T;
    if (i >= a.n)
        return;
    // Regularization<FreeSurface>: C = max(S, 1); <Internal>: C = S
    T S = a.S[i];
    T c = flow == SPHX_FLOW_FREE_SURFACE ? dmax(S, T(1)) : S;
    a.C[i] = Prec<P>::c_minus_one ? c - T(1) : c;
    a.rho[i] = c * a.rho0;
}

template <int D, int P>
__global__ void k_regularize(FluidArgs<typename Prec<P>::T> a, int flow)
{
    regularize_body<D, P>(a, blockIdx.x * blockDim.x + threadIdx.x, flow);
}

struct PosArgs
{
    double *pos[3], *dpos[3];
    uint32_t n;
};

// AdvectionStepSetup (fluid_time_step_ck.h:148-152), fused with the refresh
// of the (x, y, z, Vol) gather record
template <int D, int P>
__device__ __forceinline__ void setup_body(const FluidArgs<typename Prec<P>::T> &a, const PosArgs &pa,
                                           typename Prec<P>::T *Vol, Rec4<typename Prec<P>::T> *posvol, uint32_t i)
{
    using T = typename Prec<P>::T;
    if (i >= a.n)
        return;
    T vol = a.mass[i] / a.rho[i];
    Vol[i] = vol;
    Rec4<T> r;
    r.x = T(pa.pos[0][i]);
    r.y = T(pa.pos[1][i]);
    if (D == 3)
    {
        r.z = T(pa.pos[2][i]);
        r.w = vol;
        pa.dpos[2][i] = 0.0;
    }
    else
    {
        r.z = vol;
        r.w = T(0);
    }
    posvol[i] = r;
    pa.dpos[0][i] = 0.0;
    pa.dpos[1][i] = 0.0;
}

template <int D, int P>
__global__ void k_setup(FluidArgs<typename Prec<P>::T> a, PosArgs pa, typename Prec<P>::T *Vol,
                        Rec4<typename Prec<P>::T> *posvol)
{
    setup_body<D, P>(a, pa, Vol, posvol, blockIdx.x * blockDim.x + threadIdx.x);
}

// UpdateParticlePosition (fluid_time_step_ck.h:176-179) + record refresh
template <int D, class T>
__device__ __forceinline__ void position_body(const PosArgs &pa, const T *Vol, Rec4<T> *posvol, Rec4<float> *pos_f32,
                                              uint32_t i)
{
    if (i >= pa.n)
        return;
    Rec4<T> r;
    double x = pa.pos[0][i] + pa.dpos[0][i];
    double y = pa.pos[1][i] + pa.dpos[1][i];
    pa.pos[0][i] = x;
    pa.pos[1][i] = y;
    r.x = T(x);
    r.y = T(y);
    if (D == 3)
    {
        double z = pa.pos[2][i] + pa.dpos[2][i];
        pa.pos[2][i] = z;
        r.z = T(z);
        r.w = Vol[i];
    }
    else
    {
        r.z = Vol[i];
        r.w = T(0);
    }
    posvol[i] = r;
    if (pos_f32 != (Rec4<float> *)posvol)
        pos_f32[i] = Rec4<float>{(float)r.x, (float)r.y, D == 3 ? (float)r.z : 0.0f, 0.0f};
}

template <int D, class T>
__global__ void k_update_position(PosArgs pa, const T *Vol, Rec4<T> *posvol, Rec4<float> *pos_f32)
{
    position_body<D, T>(pa, Vol, posvol, pos_f32, blockIdx.x * blockDim.x + threadIdx.x);
}

template <int D, class T>
__global__ void k_refresh_posvol(PosArgs pa, const T *Vol, Rec4<T> *posvol, Rec4<float> *pos_f32)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pa.n)
        return;
    Rec4<T> r;
    r.x = T(pa.pos[0][i]);
    r.y = T(pa.pos[1][i]);
    if (D == 3)
    {
        r.z = T(pa.pos[2][i]);
        r.w = Vol[i];
    }
    else
    {
        r.z = Vol[i];
        r.w = T(0);
    }
    posvol[i] = r;
    if (pos_f32 != (Rec4<float> *)posvol)
        pos_f32[i] = Rec4<float>{(float)r.x, (float)r.y, D == 3 ? (float)r.z : 0.0f, 0.0f};
}

template <int D, class T>
__global__ void k_gravity(uint32_t n, T *fp0, T *fp1, T *fp2, const T *mass, double g0, double g1, double g2)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    // GravityForceCK + ForcePriorCK::update: force_prior += m g - previous(=0)
    fp0[i] += mass[i] * T(g0);
    fp1[i] += mass[i] * T(g1);
    if (D == 3)
        fp2[i] += mass[i] * T(g2);
}

//---------------------------------------------------------------------------
// reductions
//---------------------------------------------------------------------------
__device__ __forceinline__ void block_max_to(double v, unsigned long long *slot)
{
    __shared__ double red[kThreads / 64];
    double m = wave_max(v);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double b = red[0];
        for (int w = 1; w < kThreads / 64; ++w)
            b = fmax(b, red[w]);
        atomic_max_spread(slot, b);
    }
}

template <int D, int P>
__device__ __forceinline__ void signal_body(const FluidArgs<typename Prec<P>::T> &a, uint32_t blk,
                                            unsigned long long *slot)
{
    using T = typename Prec<P>::T;
    uint32_t i = blk * kThreads + threadIdx.x;
    double v = -1.0;
    if (i < a.n)
    {
        const T m_i = a.mass[i];
        T gx = a.force[0][i] + a.force_prior[0][i], gy = a.force[1][i] + a.force_prior[1][i];
        T gz = D == 3 ? a.force[2][i] + a.force_prior[2][i] : T(0);
        T fn = sqrt(D == 3 ? gx * gx + gy * gy + gz * gz : gx * gx + gy * gy);
        T acc_scale = sqrt(T(4) * a.h * fn / m_i);
        Rec4<T> vi = a.vel[i];
        T vn = sqrt(D == 3 ? vi.x * vi.x + vi.y * vi.y + vi.z * vi.z : vi.x * vi.x + vi.y * vi.y);
        v = signal_or_inf(a.c0 + vn, acc_scale, T(0));
    }
    block_max_to(v, slot);
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_signal_max(FluidArgs<typename Prec<P>::T> a,
                                                         unsigned long long *slot)
{
    signal_body<D, P>(a, blockIdx.x, slot);
}

template <int D, int P>
__device__ __forceinline__ void speed_body(const FluidArgs<typename Prec<P>::T> &a, uint32_t blk,
                                           unsigned long long *slot)
{
    using T = typename Prec<P>::T;
    uint32_t i = blk * kThreads + threadIdx.x;
    double v = -1.0;
    if (i < a.n)
    {
        Rec4<T> vi = a.vel[i];
        const T v2 = D == 3 ? vi.x * vi.x + vi.y * vi.y + vi.z * vi.z : vi.x * vi.x + vi.y * vi.y;
        v = isfinite(v2) ? (double)v2 : __builtin_huge_val();
    }
    block_max_to(v, slot);
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_speed_max(FluidArgs<typename Prec<P>::T> a,
                                                        unsigned long long *slot)
{
    speed_body<D, P>(a, blockIdx.x, slot);
}

__global__ void __launch_bounds__(kThreads) k_axis_max(const double *x, uint32_t n, unsigned long long *slot)
{
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    double v = i < n ? x[i] : -1.79769313486231570e308;
    block_max_to(v, slot);
}

// TotalMechanicalEnergyCK: sum 0.5 m |v|^2 + m g.(0 - x), fixed-order two-stage sum
template <int D, int P>
__global__ void __launch_bounds__(kThreads) k_energy_parts(FluidArgs<typename Prec<P>::T> a, PosArgs pa,
                                                           double g0, double g1, double g2, double *parts)
{
    using T = typename Prec<P>::T;
    __shared__ double red[kThreads / 64];
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    double e = 0.0;
    if (i < a.n)
    {
        Rec4<T> vi = a.vel[i];
        T m = a.mass[i];
        T v2 = D == 3 ? vi.x * vi.x + vi.y * vi.y + vi.z * vi.z : vi.x * vi.x + vi.y * vi.y;
        double pot = g0 * (0.0 - pa.pos[0][i]) + g1 * (0.0 - pa.pos[1][i]);
        if (D == 3)
            pot += g2 * (0.0 - pa.pos[2][i]);
        e = (double)(T(0.5) * m * v2) + (double)m * pot;
    }
    double s = wave_sum(e);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double b = 0.0;
        for (int w = 0; w < kThreads / 64; ++w)
            b += red[w];
        parts[blockIdx.x] = b;
    }
}

__global__ void __launch_bounds__(kThreads) k_sum_parts(const double *parts, uint32_t nparts, double *out)
{
    __shared__ double red[kThreads / 64];
    double s = 0.0;
    for (uint32_t k = threadIdx.x; k < nparts; k += kThreads)
        s += parts[k];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double b = 0.0;
        for (int w = 0; w < kThreads / 64; ++w)
            b += red[w];
        *out = b;
    }
}

__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Interpolation<Contact<Real>>::interact (interpolation_dynamics.hpp:46-60)
template <int D, class T>
__global__ void k_interpolate(const Rec4<T> *obs, uint32_t n_obs, const Rec4<T> *tar, const T *data,
                              uint32_t stride, const uint32_t *chunk, const uint32_t *count, const uint32_t *slice,
                              const KernelPoly<T> *poly, T *out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obs)
        return;
    const PV<T, D> pi = load_pv<T, D>(obs, i);
    T q = 0, w = 0;
    const uint32_t nb = chunk[i >> 6] + (i & 63u), nc = count[i];
    for (uint32_t k = 0; k < nc; ++k)
    {
        const uint32_t j = slice[nb + (k << 6)];
        const PV<T, D> pj = load_pv<T, D>(tar, j);
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = D == 3 ? pi.z - pj.z : T(0);
        const T r2 = D == 3 ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
        const T x = sqrt(r2) * poly->x_per_r;
        const int loc = min((int)x, kPolyIntervals - 1);
        const T f = frac_part(x);
        const T *c = poly->w[loc];
        const T W = ((c[3] * f + c[2]) * f + c[1]) * f + c[0];
        const T weight = W * pj.vol;
        q += weight * data[(size_t)j * stride];
        w += weight;
    }
    out[i] = q / (w + T(2.71051e-20));
}

//---------------------------------------------------------------------------
// device-side dual-time loop of sphx_advection_step (dambreak.cpp:188-207):
// AdvectionTimeStepCK's dt from the speed max, then per sub-step
// AcousticTimeStepCK's dt from the signal max and the loop condition
// relaxation < advection dt, with the host's arithmetic in double
//---------------------------------------------------------------------------
// the max of the accumulator lanes, which are then reset for the next reduction
__device__ __forceinline__ double lanes_max_reset(unsigned long long *lanes)
{
    unsigned long long m = 0;
    for (int k = 0; k < kMaxLanes; ++k)
    {
        m = m > lanes[k * kMaxStride] ? m : lanes[k * kMaxStride];
        lanes[k * kMaxStride] = kOrderedLowest;
    }
    return from_ordered_bits(m);
}
__device__ __forceinline__ void put_double(unsigned long long *slot, double v)
{
    *slot = (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double get_double(const unsigned long long *slot)
{
    return __longlong_as_double((long long)*slot);
}

__device__ __forceinline__ void step_begin_body(unsigned long long *scalars, double u_ref, double cfl, double h,
                                                double time)
{
    unsigned long long *st = scalars + SLOT_STEP;
    const double m = lanes_max_reset(scalars + SLOT_MAX);
    // AdvectionTimeStepCK::FinishDynamics::Result, fluid_time_step_ck.cpp:23-26
    const double adv_dt = cfl * h / (fmax(sqrt(m), u_ref) + 2.71051e-20);
    put_double(st + STEP_ADV_DT, adv_dt);
    put_double(st + STEP_DT, 0.0);
    put_double(st + STEP_RELAXATION, 0.0);
    put_double(st + STEP_TIME, time);
    st[STEP_COUNT] = 0;
    st[STEP_ACTIVE] = 0;
    // a non-finite velocity made the speed max +inf and the step zero
    st[STEP_ERROR] = isfinite(adv_dt) && adv_dt > 0.0 ? 0ull : 1ull;
}

__global__ void k_step_begin(unsigned long long *scalars, double u_ref, double cfl, double h, double time)
{
    if (threadIdx.x == 0)
        step_begin_body(scalars, u_ref, cfl, h, time);
}

__device__ __forceinline__ void substep_begin_body(unsigned long long *scalars, double cfl, double h)
{
    unsigned long long *st = scalars + SLOT_STEP;
    const double relaxation = get_double(st + STEP_RELAXATION);
    if (st[STEP_ERROR] != 0 || !(relaxation < get_double(st + STEP_ADV_DT)))
    {
        st[STEP_ACTIVE] = 0;
        return;
    }
    // AcousticTimeStepCK::FinishDynamics::Result, fluid_time_step_ck.hpp:30-57
    const double dt = cfl * h / (lanes_max_reset(scalars + SLOT_MAX) + 2.71051e-20);
    if (!isfinite(dt) || !(dt > 0.0))
    {
        st[STEP_ERROR] = 1;
        st[STEP_ACTIVE] = 0;
        return;
    }
    put_double(st + STEP_DT, dt);
    put_double(st + STEP_RELAXATION, relaxation + dt);
    put_double(st + STEP_TIME, get_double(st + STEP_TIME) + dt);
    st[STEP_COUNT] += 1;
    st[STEP_ACTIVE] = 1;
}

__global__ void k_substep_begin(unsigned long long *scalars, double cfl, double h)
{
    if (threadIdx.x == 0)
        substep_begin_body(scalars, cfl, h);
}

//---------------------------------------------------------------------------
// Batched launches: one grid over the blocks of every scene of a batch.
// map[b] = scene << 16 | block of that scene, laid out scene after scene, so
// that the XCD-aware block order keeps each scene's blocks on one XCD where
// it can; every block reads its scene's arguments from the scene table.
// The tail kernels of an advection step run only for the scenes whose gate
// is set (their sub-steps are done and the tail is not left to the host).
//---------------------------------------------------------------------------
template <class T>
struct SceneArgs
{
    FluidArgs<T> f;
    PosArgs pa;
    T *Vol;
    Rec4<T> *posvol;
    Rec4<float> *pos_f32;
    unsigned long long *scalars; // the scene context's scalar slots (step state, max lanes)
    const double *time;          // physical time at the start of the step
    double u_ref, advection_cfl, acoustic_cfl, h;
    int flow;
};

__device__ __forceinline__ uint32_t map_scene(uint32_t m) { return m >> 16; }
__device__ __forceinline__ uint32_t map_block(uint32_t m) { return m & 0xffffu; }

template <int D, int P>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(P == P_MIXED ? 6 : 1)))
kb_summation(const SceneArgs<typename Prec<P>::T> *__restrict__ tab, const uint32_t *__restrict__ map)
{
    const uint32_t m = map[xcd_block_index()];
    summation_body<D, P>(tab[map_scene(m)].f, map_block(m));
}

// DensityRegularization + AdvectionStepSetup, which touch each particle alone
template <int D, int P>
__global__ void __launch_bounds__(kThreads) kb_regularize_setup(const SceneArgs<typename Prec<P>::T> *__restrict__ tab,
                                                                const uint32_t *__restrict__ map)
{
    const uint32_t m = map[blockIdx.x];
    const SceneArgs<typename Prec<P>::T> &sa = tab[map_scene(m)];
    const uint32_t i = map_block(m) * kThreads + threadIdx.x;
    regularize_body<D, P>(sa.f, i, sa.flow);
    setup_body<D, P>(sa.f, sa.pa, sa.Vol, sa.posvol, i);
}

// the max lanes of every scene to the lowest value (one block per scene)
template <class T>
__global__ void kb_reset_lanes(const SceneArgs<T> *__restrict__ tab, uint32_t n_scenes)
{
    if (blockIdx.x < n_scenes && threadIdx.x < (unsigned)kMaxLanes)
        tab[blockIdx.x].scalars[SLOT_MAX + threadIdx.x * kMaxStride] = kOrderedLowest;
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) kb_speed_max(const SceneArgs<typename Prec<P>::T> *__restrict__ tab,
                                                         const uint32_t *__restrict__ map)
{
    const uint32_t m = map[blockIdx.x];
    const SceneArgs<typename Prec<P>::T> &sa = tab[map_scene(m)];
    speed_body<D, P>(sa.f, map_block(m), sa.scalars + SLOT_MAX);
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) kb_signal_max(const SceneArgs<typename Prec<P>::T> *__restrict__ tab,
                                                          const uint32_t *__restrict__ map)
{
    const uint32_t m = map[blockIdx.x];
    const SceneArgs<typename Prec<P>::T> &sa = tab[map_scene(m)];
    signal_body<D, P>(sa.f, map_block(m), sa.scalars + SLOT_MAX);
}

// one thread per scene
template <class T>
__global__ void kb_step_begin(const SceneArgs<T> *__restrict__ tab, uint32_t n_scenes)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_scenes)
    {
        const SceneArgs<T> &sa = tab[s];
        step_begin_body(sa.scalars, sa.u_ref, sa.advection_cfl, sa.h, *sa.time);
    }
}
template <class T>
__global__ void kb_substep_begin(const SceneArgs<T> *__restrict__ tab, uint32_t n_scenes)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_scenes)
        substep_begin_body(tab[s].scalars, tab[s].acoustic_cfl, tab[s].h);
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) kb_init_1st(const SceneArgs<typename Prec<P>::T> *__restrict__ tab,
                                                        const uint32_t *__restrict__ map)
{
    const uint32_t m = map[blockIdx.x];
    init_1st_body<D, P>(tab[map_scene(m)].f, map_block(m), 0.0);
}

template <int D, int P, bool RIEMANN>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(P == P_MIXED ? 5 : 1))) kb_pass1(const SceneArgs<typename Prec<P>::T> *__restrict__ tab,
                                                     const uint32_t *__restrict__ map)
{
    const uint32_t m = map[xcd_block_index()];
    pass1_body<D, P, RIEMANN, true, false>(tab[map_scene(m)].f, map_block(m), 0.0);
}

template <int D, int P, bool RIEMANN>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(P == P_MIXED ? 4 : 1)))
kb_pass2(const SceneArgs<typename Prec<P>::T> *__restrict__ tab, const uint32_t *__restrict__ map)
{
    const uint32_t m = map[xcd_block_index()];
    const SceneArgs<typename Prec<P>::T> &sa = tab[map_scene(m)];
    pass2_body<D, P, RIEMANN, true, false>(sa.f, map_block(m), 0.0, sa.scalars + SLOT_MAX);
}

// after the captured sub-steps: gate[s] = 1 when scene s finished its loop
// without error and its tail is run on the device (tail_on[s]); its step
// state is copied to state[s * STEP_SLOTS ...] for the host
template <class T>
__global__ void kb_gate(const SceneArgs<T> *__restrict__ tab, uint32_t n_scenes, const uint32_t *__restrict__ tail_on,
                        uint32_t *gate, unsigned long long *state)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_scenes)
        return;
    const unsigned long long *st = tab[s].scalars + SLOT_STEP;
    for (int k = 0; k < STEP_SLOTS; ++k)
        state[(size_t)s * STEP_SLOTS + k] = st[k];
    const bool done = st[STEP_ERROR] == 0 && !(get_double(st + STEP_RELAXATION) < get_double(st + STEP_ADV_DT));
    gate[s] = done && tail_on[s] ? 1u : 0u;
}

template <int D, int P>
__global__ void __launch_bounds__(kThreads) kb_update_position(const SceneArgs<typename Prec<P>::T> *__restrict__ tab,
                                                               const uint32_t *__restrict__ map,
                                                               const uint32_t *__restrict__ gate)
{
    const uint32_t m = map[blockIdx.x];
    const uint32_t sc = map_scene(m);
    if (!gate[sc])
        return;
    const SceneArgs<typename Prec<P>::T> &sa = tab[sc];
    position_body<D, typename Prec<P>::T>(sa.pa, sa.Vol, sa.posvol, sa.pos_f32, map_block(m) * kThreads + threadIdx.x);
}

} // namespace

//=============================================================================
// host side
//=============================================================================
template <class T>
static FluidArgs<T> fluid_args(sphx_body *b)
{
    FluidArgs<T> a{};
    a.n = b->n;
    a.step = b->ctx->device_step ? b->ctx->scalars_d + SLOT_STEP : nullptr;
    a.posvol = (const Rec4<T> *)b->posvol;
    a.vel = (Rec4<T> *)b->vel;
    for (int k = 0; k < 3; ++k)
    {
        a.force[k] = (T *)b->force[k];
        a.force_prior[k] = (T *)b->force_prior[k];
        a.dpos[k] = b->dpos[k];
    }
    a.mass = (const T *)b->mass;
    a.rho = (T *)b->rho;
    a.p = (T *)b->p;
    a.C = (T *)b->C;
    a.dC = (T *)b->dC;
    a.S = (T *)b->S;
    a.Vol_ref = (const T *)b->Vol_ref;
    for (int c = 0; c < 9; ++c)
        a.corr[c] = (T *)b->corr[c];
    for (int k = 0; k < 3; ++k)
        a.grad_sum[k] = (T *)b->grad_sum[k];
    a.indicator = b->indicator;
    a.prev_indicator = b->prev_indicator;
    a.pos_div = (T *)b->pos_div;
    for (int k = 0; k < 3; ++k)
    {
        a.visc[k] = (T *)b->visc[k];
        a.prev_visc[k] = (T *)b->prev_visc[k];
        a.kgi[k] = (T *)b->kgi[k];
    }
    const double rho0 = b->desc.rho0, c0 = b->desc.c0;
    a.rho0 = T(rho0);
    a.c0 = T(c0);
    a.p0 = T(rho0 * c0 * c0);
    a.h = T(b->h);
    // ImpedanceModel for fluid_i == fluid_j (riemann_solver_ck.hpp:59-69), in double
    const double z = rho0 * c0;
    const double inv_sum = 1.0 / (z + z);
    const double inv_ave = (z + z) / (z * z + z * z);
    a.rho0c0 = T(z);
    a.inv_rho0c0_sum = T(inv_sum);
    a.inv_rho0c0_ave = T(inv_ave);
    a.rho0c0_geo_ave = T(2.0 * z * z * inv_sum);
    a.inv_c0_ave = T(0.5 * (rho0 + rho0) * inv_ave);
    return a;
}

template <class T>
static void attach_relations(FluidArgs<T> &a, sphx_relation *inner, sphx_relation *const *w, int nw)
{
    if (inner)
    {
        a.chunk = inner->chunk_offset;
        a.count = inner->neighbor_size;
        a.slice = inner->slice;
        if (inner->staged && !inner->stage_stale)
        {
            a.stage_count = inner->stage_count;
            a.stage_keys = inner->stage_keys;
            a.local_slice = inner->local_slice;
            a.stage_cap = inner->stage_cap;
        }
        a.poly = (const KernelPoly<T> *)inner->poly;
    }
    a.n_wall = nw;
    for (int k = 0; k < nw; ++k)
    {
        sphx_body *wb = w[k]->target;
        WallArgs<T> &x = a.wall[k];
        x.posvol = (const Rec4<T> *)wb->posvol;
        for (int d = 0; d < 3; ++d)
        {
            x.vel[d] = (const T *)wb->wall_vel[d];
            x.acc[d] = (const T *)wb->wall_acc[d];
            x.normal[d] = (const T *)wb->normal[d];
        }
        x.Vol_ref = (const T *)wb->Vol_ref;
        x.chunk = w[k]->chunk_offset;
        x.count = w[k]->neighbor_size;
        x.slice = w[k]->slice;
        x.poly = (const KernelPoly<T> *)w[k]->poly;
    }
}

static PosArgs pos_args(sphx_body *b)
{
    PosArgs p{};
    for (int k = 0; k < 3; ++k)
    {
        p.pos[k] = b->pos[k];
        p.dpos[k] = b->dpos[k];
    }
    p.n = b->n;
    return p;
}

// dispatch a templated launcher over (D, P)
template <template <int, int> class L, class... A>
static sphx_status dispatch(sphx_body *b, A &&...args)
{
    if (b->D == 2 && b->P == P_FP64)
        return L<2, P_FP64>::run(b, args...);
    if (b->D == 3 && b->P == P_FP64)
        return L<3, P_FP64>::run(b, args...);
    if (b->D == 2 && b->P == P_MIXED)
        return L<2, P_MIXED>::run(b, args...);
    return L<3, P_MIXED>::run(b, args...);
}

// kernels that take more than 64 KiB of dynamic LDS must say so first
template <class K>
static void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

#define LAUNCH_CHECK()                     \
    do                                     \
    {                                      \
        SPHX_HIP(hipGetLastError());       \
    } while (0)

template <int D, int P>
struct RefreshL
{
    static sphx_status run(sphx_body *b)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        hipLaunchKernelGGL((k_refresh_posvol<D, T>), dim3(grid_for(b->n, 256)), dim3(256), 0, b->ctx->stream,
                           pos_args(b), (const T *)b->Vol, (Rec4<T> *)b->posvol, b->pos_f32);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status refresh_posvol(sphx_body *b)
{
    b->state_version++;
    return dispatch<RefreshL>(b);
}

template <int D, int P>
struct GravityL
{
    static sphx_status run(sphx_body *b, const double *g)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        hipLaunchKernelGGL((k_gravity<D, T>), dim3(grid_for(b->n, 256)), dim3(256), 0, b->ctx->stream, b->n,
                           (T *)b->force_prior[0], (T *)b->force_prior[1], (T *)b->force_prior[2],
                           (const T *)b->mass, g[0], g[1], D == 3 ? g[2] : 0.0);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status gravity_force(sphx_body *b, const double *g) { return dispatch<GravityL>(b, g); }

template <int D, int P>
struct RegularizeL
{
    static sphx_status run(sphx_body *b, int flow)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        ProfileScope prof(b->ctx, "state_update");
        hipLaunchKernelGGL((k_regularize<D, P>), dim3(grid_for(b->n, 256)), dim3(256), 0, b->ctx->stream,
                           fluid_args<T>(b), flow);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status density_regularization(sphx_body *b, int flow) { return dispatch<RegularizeL>(b, flow); }

template <int D, int P>
struct SetupL
{
    static sphx_status run(sphx_body *b)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        ProfileScope prof(b->ctx, "state_update");
        hipLaunchKernelGGL((k_setup<D, P>), dim3(grid_for(b->n, 256)), dim3(256), 0, b->ctx->stream,
                           fluid_args<T>(b), pos_args(b), (T *)b->Vol, (Rec4<T> *)b->posvol);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status advection_step_setup(sphx_body *b)
{
    b->state_version++;
    return dispatch<SetupL>(b);
}

template <int D, int P>
struct PositionL
{
    static sphx_status run(sphx_body *b)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        ProfileScope prof(b->ctx, "state_update");
        hipLaunchKernelGGL((k_update_position<D, T>), dim3(grid_for(b->n, 256)), dim3(256), 0, b->ctx->stream,
                           pos_args(b), (const T *)b->Vol, (Rec4<T> *)b->posvol, b->pos_f32);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status update_position(sphx_body *b)
{
    b->state_version++;
    return dispatch<PositionL>(b);
}

template <int D, int P>
struct SummationL
{
    static sphx_status run(sphx_body *b, sphx_relation *inner, sphx_relation *const *c, int nc)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        attach_relations(a, inner, c, nc);
        ProfileScope prof(b->ctx, "compression_summation");
        const size_t lds = a.stage_count ? (size_t)a.stage_cap * sizeof(Rec4<T>) : 0;
        allow_lds(k_summation<D, P>, lds);
        hipLaunchKernelGGL((k_summation<D, P>), dim3(grid_for(b->n, kThreads)), dim3(kThreads), lds,
                           b->ctx->stream, a);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status compression_summation(sphx_relation *inner, sphx_relation *const *c, int nc)
{
    return dispatch<SummationL>(inner->source, inner, c, nc);
}

// launch one of the eight boolean instantiations of a pass kernel
#define SPHX_PASS_VARIANTS(K, B0, B1, B2, ...)                                                                   \
    do                                                                                                           \
    {                                                                                                            \
        if (B0 && B1 && B2)                                                                                      \
            LAUNCH_PASS((K<D, P, true, true, true>), __VA_ARGS__);                                               \
        else if (B0 && B1)                                                                                       \
            LAUNCH_PASS((K<D, P, true, true, false>), __VA_ARGS__);                                              \
        else if (B0 && B2)                                                                                       \
            LAUNCH_PASS((K<D, P, true, false, true>), __VA_ARGS__);                                              \
        else if (B0)                                                                                             \
            LAUNCH_PASS((K<D, P, true, false, false>), __VA_ARGS__);                                             \
        else if (B1 && B2)                                                                                       \
            LAUNCH_PASS((K<D, P, false, true, true>), __VA_ARGS__);                                              \
        else if (B1)                                                                                             \
            LAUNCH_PASS((K<D, P, false, true, false>), __VA_ARGS__);                                             \
        else if (B2)                                                                                             \
            LAUNCH_PASS((K<D, P, false, false, true>), __VA_ARGS__);                                             \
        else                                                                                                     \
            LAUNCH_PASS((K<D, P, false, false, false>), __VA_ARGS__);                                            \
    } while (0)
#define LAUNCH_PASS(KERNEL, ...)                                                                                 \
    do                                                                                                           \
    {                                                                                                            \
        allow_lds(KERNEL, lds);                                                                                  \
        hipLaunchKernelGGL(KERNEL, g, blk, lds, b->ctx->stream, __VA_ARGS__);                                   \
    } while (0)

template <int D, int P>
struct CorrectionGradL
{
    static sphx_status run(sphx_body *b, sphx_relation *inner)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        attach_relations(a, inner, nullptr, 0);
        hipLaunchKernelGGL((k_correction_grad<D, P>), dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0,
                           b->ctx->stream, a);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

// grad_sum for the current positions, volumes, correction matrices and inner lists
static sphx_status refresh_grad_sum(sphx_body *b, sphx_relation *inner)
{
    sphx_body::GradKey key{b->state_version, inner, inner->builds};
    if (b->grad_key.state == key.state && b->grad_key.rel == key.rel && b->grad_key.builds == key.builds)
        return SPHX_OK;
    ProfileScope prof(b->ctx, "linear_correction");
    sphx_status s = dispatch<CorrectionGradL>(b, inner);
    if (!s)
        b->grad_key = key;
    return s;
}

template <int D, int P>
struct Acoustic1L
{
    static sphx_status run(sphx_body *b, sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                           double dt, bool fuse_init_2nd, bool corr)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        a.riemann = riemann;
        attach_relations(a, inner, w, nw);
        dim3 g(grid_for(b->n, kThreads)), blk(kThreads);
        {
            ProfileScope prof(b->ctx, "acoustic_init");
            hipLaunchKernelGGL((k_init_1st<D, P>), g, blk, 0, b->ctx->stream, a, dt);
        }
        ProfileScope prof(b->ctx, "acoustic_pass1");
        const bool rs = riemann != SPHX_RIEMANN_NONE;
        const size_t lds = a.stage_count ? (size_t)pass1_cap<P>() * (sizeof(Rec4<T>) + sizeof(T)) : 0;
        SPHX_PASS_VARIANTS(k_pass1, rs, fuse_init_2nd, corr, a, dt);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

template <int D, int P>
struct Acoustic2L
{
    static sphx_status run(sphx_body *b, sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                           double dt, bool run_init, unsigned long long *max_slot, bool corr)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        a.riemann = riemann;
        attach_relations(a, inner, w, nw);
        dim3 g(grid_for(b->n, kThreads)), blk(kThreads);
        if (run_init)
        {
            ProfileScope prof(b->ctx, "acoustic_init");
            hipLaunchKernelGGL((k_init_2nd<D, P>), g, blk, 0, b->ctx->stream, a, dt);
        }
        ProfileScope prof(b->ctx, "acoustic_pass2");
        const bool rs = riemann != SPHX_RIEMANN_NONE;
        const bool cfl = max_slot != nullptr;
        const size_t lds = a.stage_count ? (size_t)kPass2Cap * (sizeof(Rec4<T>) + 3 * sizeof(T)) : 0;
        SPHX_PASS_VARIANTS(k_pass2, rs, cfl, corr, a, dt, max_slot);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

sphx_status acoustic_1st_half(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann, double dt,
                              bool fuse_init_2nd)
{
    return dispatch<Acoustic1L>(inner->source, inner, w, nw, riemann, dt, fuse_init_2nd, false);
}

sphx_status acoustic_2nd_half(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann, double dt,
                              bool run_init, bool fuse_cfl)
{
    sphx_body *b = inner->source;
    unsigned long long *slot = fuse_cfl ? b->ctx->scalars_d + SLOT_MAX : nullptr;
    return dispatch<Acoustic2L>(b, inner, w, nw, riemann, dt, run_init, slot, false);
}

sphx_status acoustic_2nd_half_traversal(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann, double dt)
{
    sphx_body *b = inner->source;
    if (b->D != 3 || b->P != P_MIXED)
    {
        set_error("the cell traversal runs 3D FP32-mixed bodies");
        return SPHX_E_ARG;
    }
    if (!b->n)
        return SPHX_OK;
    const sphx_cell_linked_list *c = inner->cll;
    FluidArgs<float> a = fluid_args<float>(b);
    a.riemann = riemann;
    attach_relations(a, inner, w, nw);
    a.stage_count = nullptr; // the traversal stages its own boxes
    TravArgs t;
    t.cell_offset = c->cell_offset;
    t.particle_index = c->particle_index;
    t.m = c->mesh;
    t.pos = PosPtrs{b->pos[0], b->pos[1], b->pos[2]};
    t.inv_h = inner->inv_h;
    relation_prefilter(inner, &t.thr_lo, &t.thr_hi, &t.pre_xmax);
    {
        ProfileScope prof(b->ctx, "acoustic_init");
        hipLaunchKernelGGL((k_init_2nd<3, P_MIXED>), dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0,
                           b->ctx->stream, a, dt);
    }
    ProfileScope prof(b->ctx, "acoustic_pass2");
    const size_t lds = (size_t)kTravCap * (sizeof(Rec4<float>) + sizeof(uint32_t));
    const dim3 g(grid_for(b->n, 64)), blk(64);
    if (riemann != SPHX_RIEMANN_NONE)
    {
        allow_lds(k_pass2_traverse<true>, lds);
        hipLaunchKernelGGL(k_pass2_traverse<true>, g, blk, lds, b->ctx->stream, a, t, dt);
    }
    else
    {
        allow_lds(k_pass2_traverse<false>, lds);
        hipLaunchKernelGGL(k_pass2_traverse<false>, g, blk, lds, b->ctx->stream, a, t, dt);
    }
    LAUNCH_CHECK();
    return SPHX_OK;
}

sphx_status acoustic_1st_half_corrected(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                                        double dt, bool fuse_init_2nd)
{
    sphx_status s = refresh_grad_sum(inner->source, inner);
    if (s)
        return s;
    return dispatch<Acoustic1L>(inner->source, inner, w, nw, riemann, dt, fuse_init_2nd, true);
}

sphx_status acoustic_2nd_half_corrected(sphx_relation *inner, sphx_relation *const *w, int nw, int riemann,
                                        double dt, bool run_init, bool fuse_cfl)
{
    sphx_body *b = inner->source;
    unsigned long long *slot = fuse_cfl ? b->ctx->scalars_d + SLOT_MAX : nullptr;
    return dispatch<Acoustic2L>(b, inner, w, nw, riemann, dt, run_init, slot, true);
}

template <int D, int P>
struct CorrectionL
{
    static sphx_status run(sphx_body *b, sphx_relation *inner, sphx_relation *const *c, int nc, double alpha)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        attach_relations(a, inner, c, nc);
        const size_t lds = a.stage_count ? (size_t)a.stage_cap * sizeof(Rec4<T>) : 0;
        allow_lds(k_correction<D, P>, lds);
        hipLaunchKernelGGL((k_correction<D, P>), dim3(grid_for(b->n, kThreads)), dim3(kThreads), lds,
                           b->ctx->stream, a, alpha);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

sphx_status linear_correction_matrix(sphx_relation *inner, sphx_relation *const *c, int nc, double alpha)
{
    sphx_body *b = inner->source;
    ProfileScope prof(b->ctx, "linear_correction");
    b->state_version++;
    return dispatch<CorrectionL>(b, inner, c, nc, alpha);
}

template <int D, int P>
struct ViscousL
{
    static sphx_status run(sphx_body *b, sphx_relation *inner, sphx_relation *const *w, int nw, double mu,
                           bool corr)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        attach_relations(a, inner, w, nw);
        // Viscosity::getInterParticleViscosity: PairGeomAverageFixed(mu, mu), particle_functors.h:169-170
        const double mu_pair = 2.0 * mu * mu * (1.0 / (mu + mu));
        const double h_sq = std::pow(b->h, 2.0);
        dim3 g(grid_for(b->n, kThreads)), blk(kThreads);
        if (corr)
            hipLaunchKernelGGL((k_viscous<D, P, true>), g, blk, 0, b->ctx->stream, a, mu_pair, mu, h_sq);
        else
            hipLaunchKernelGGL((k_viscous<D, P, false>), g, blk, 0, b->ctx->stream, a, mu_pair, mu, h_sq);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

sphx_status viscous_force(sphx_relation *inner, sphx_relation *const *w, int nw, double mu, bool corrected)
{
    ProfileScope prof(inner->source->ctx, "viscous_force");
    return dispatch<ViscousL>(inner->source, inner, w, nw, mu, corrected);
}

template <int D, int P>
struct KgiL
{
    static sphx_status run(sphx_body *b, sphx_relation *inner, sphx_relation *const *c, int nc, bool corr)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        attach_relations(a, inner, c, nc);
        dim3 g(grid_for(b->n, kThreads)), blk(kThreads);
        if (corr)
            hipLaunchKernelGGL((k_kernel_gradient_integral<D, P, true>), g, blk, 0, b->ctx->stream, a);
        else
            hipLaunchKernelGGL((k_kernel_gradient_integral<D, P, false>), g, blk, 0, b->ctx->stream, a);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

sphx_status kernel_gradient_integral(sphx_relation *inner, sphx_relation *const *c, int nc, bool corrected)
{
    ProfileScope prof(inner->source->ctx, "transport_correction");
    return dispatch<KgiL>(inner->source, inner, c, nc, corrected);
}

template <int D, int P>
struct TransportL
{
    static sphx_status run(sphx_body *b, double coefficient, int limiter, bool bulk_only)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        const double sq_h = b->h * b->h;
        hipLaunchKernelGGL((k_transport_correction<D, P>), dim3(grid_for(b->n, 256)), dim3(256), 0, b->ctx->stream,
                           fluid_args<T>(b), coefficient * sq_h, sq_h,
                           limiter == SPHX_LIMITER_TRUNCATED_LINEAR ? 1 : 0, bulk_only ? 1 : 0);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

sphx_status transport_velocity_correction(sphx_body *b, double coefficient, int limiter, bool bulk_only)
{
    ProfileScope prof(b->ctx, "transport_correction");
    return dispatch<TransportL>(b, coefficient, limiter, bulk_only);
}

template <int D, int P>
struct SurfaceL
{
    static sphx_status run(sphx_body *b, sphx_relation *inner, sphx_relation *const *c, int nc)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        FluidArgs<T> a = fluid_args<T>(b);
        attach_relations(a, inner, c, nc);
        dim3 g(grid_for(b->n, kThreads)), blk(kThreads);
        hipLaunchKernelGGL((k_surface_div<D, P>), g, blk, 0, b->ctx->stream, a);
        hipLaunchKernelGGL((k_surface_update<D, P>), g, blk, 0, b->ctx->stream, a);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};

sphx_status free_surface_indication(sphx_relation *inner, sphx_relation *const *c, int nc)
{
    ProfileScope prof(inner->source->ctx, "surface_indication");
    return dispatch<SurfaceL>(inner->source, inner, c, nc);
}

sphx_status fill_u64(sphx_context *ctx, unsigned long long *p, size_t n, unsigned long long v)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 64)), dim3(64), 0, ctx->stream, p, n, v);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

template <int D, int P>
struct SignalL
{
    static sphx_status run(sphx_body *b, unsigned long long *slot)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        ProfileScope prof(b->ctx, "time_step_reduce");
        hipLaunchKernelGGL((k_signal_max<D, P>), dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0,
                           b->ctx->stream, fluid_args<T>(b), slot);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status acoustic_signal_max(sphx_body *b, unsigned long long *slot) { return dispatch<SignalL>(b, slot); }

template <int D, int P>
struct SpeedL
{
    static sphx_status run(sphx_body *b, unsigned long long *slot)
    {
        using T = typename Prec<P>::T;
        if (!b->n)
            return SPHX_OK;
        ProfileScope prof(b->ctx, "time_step_reduce");
        hipLaunchKernelGGL((k_speed_max<D, P>), dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0,
                           b->ctx->stream, fluid_args<T>(b), slot);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status advection_speed_max(sphx_body *b, unsigned long long *slot) { return dispatch<SpeedL>(b, slot); }

sphx_status axis_max(sphx_body *b, int axis, unsigned long long *slot)
{
    if (!b->n)
        return SPHX_OK;
    hipLaunchKernelGGL(k_axis_max, dim3(grid_for(b->n, kThreads)), dim3(kThreads), 0, b->ctx->stream,
                       b->pos[axis], b->n, slot);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

template <int D, int P>
struct EnergyL
{
    static sphx_status run(sphx_body *b, const double *g, double *out)
    {
        using T = typename Prec<P>::T;
        uint32_t nb = grid_for(b->n, kThreads);
        sphx_status s = b->ctx->reduce_parts.reserve((size_t)std::max(1u, nb) * sizeof(double));
        if (s)
            return s;
        double *parts = (double *)b->ctx->reduce_parts.ptr;
        if (b->n)
            hipLaunchKernelGGL((k_energy_parts<D, P>), dim3(nb), dim3(kThreads), 0, b->ctx->stream, fluid_args<T>(b),
                               pos_args(b), g[0], g[1], D == 3 ? g[2] : 0.0, parts);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kThreads), 0, b->ctx->stream, parts, b->n ? nb : 0u, out);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status mechanical_energy(sphx_body *b, const double *g, double *out) { return dispatch<EnergyL>(b, g, out); }

template <int D, int P>
struct InterpL
{
    static sphx_status run(sphx_body *o, sphx_relation *rel, int field, int component)
    {
        using T = typename Prec<P>::T;
        if (!o->n)
            return SPHX_OK;
        sphx_body *t = rel->target;
        // a velocity component k is every 4th value of the (x, y, z, .) records from k on
        const T *data = field == SPHX_FIELD_PRESSURE   ? (const T *)t->p
                        : field == SPHX_FIELD_DENSITY  ? (const T *)t->rho
                        : field == SPHX_FIELD_VELOCITY ? (const T *)t->vel + component
                                                       : (const T *)t->C;
        const uint32_t stride = field == SPHX_FIELD_VELOCITY ? 4u : 1u;
        hipLaunchKernelGGL((k_interpolate<D, T>), dim3(grid_for(o->n, 64)), dim3(64), 0, o->ctx->stream,
                           (const Rec4<T> *)o->posvol, o->n, (const Rec4<T> *)t->posvol, data, stride,
                           rel->chunk_offset, rel->neighbor_size, rel->slice, (const KernelPoly<T> *)rel->poly,
                           (T *)o->interp);
        LAUNCH_CHECK();
        return SPHX_OK;
    }
};
sphx_status interpolate_contact(sphx_relation *rel, int field, int component)
{
    return dispatch<InterpL>(rel->source, rel, field, component);
}

//---------------------------------------------------------------------------
// kernel table -> per-interval cubic (host, long double)
// reference table: kernel_tabulated_ck.cpp:6-25, Wendland C2 kernel_wendland_c2.cpp:8-60
//---------------------------------------------------------------------------
template <class T>
static void expand_table(const double *data, long double factor, T (*out)[4], double dq)
{
    const long double d = dq;
    const long double dl0 = (-1.0 * dq) * (-2.0 * dq) * (-3.0 * dq);
    const long double dl1 = dq * (-1.0 * dq) * (-2.0 * dq);
    const long double dl2 = (2.0 * dq) * dq * (-1.0 * dq);
    const long double dl3 = (3.0 * dq) * (2.0 * dq) * dq;
    for (int loc = 0; loc < kPolyIntervals; ++loc)
    {
        const int i = loc + 1;
        // basis polynomials in t (t = fraction_1): roots as in interpolateCubic
        // L0 = t (t-d)(t-2d)/dl0, L1 = (t+d)(t-d)(t-2d)/dl1, L2 = (t+d) t (t-2d)/dl2, L3 = (t+d) t (t-d)/dl3
        auto cubic = [](long double r0, long double r1, long double r2, long double *c) {
            // (t - r0)(t - r1)(t - r2) = t^3 - (r0+r1+r2) t^2 + (r0 r1 + r0 r2 + r1 r2) t - r0 r1 r2
            c[3] = 1.0L;
            c[2] = -(r0 + r1 + r2);
            c[1] = r0 * r1 + r0 * r2 + r1 * r2;
            c[0] = -r0 * r1 * r2;
        };
        long double b0[4], b1[4], b2[4], b3[4];
        cubic(0.0L, d, 2 * d, b0);
        cubic(-d, d, 2 * d, b1);
        cubic(-d, 0.0L, 2 * d, b2);
        cubic(-d, 0.0L, d, b3);
        // coefficient of t^k times dq^k: the polynomial in f = t / dq
        long double dqk = 1.0L;
        for (int k = 0; k < 4; ++k)
        {
            long double v = b0[k] / dl0 * data[i - 1] + b1[k] / dl1 * data[i] + b2[k] / dl2 * data[i + 1] +
                            b3[k] / dl3 * data[i + 2];
            out[loc][k] = T(v * dqk * factor);
            dqk *= d;
        }
    }
}

template <class T>
static sphx_status build_poly_t(sphx_relation *rel)
{
    sphx_body *src = rel->source;
    const int D = src->D;
    // kernel of the source body (h of the source adaptation), base_kernel.cpp:6-27
    const double h = src->h;
    const double inv_h_k = 1.0 / h;
    const double Pi = M_PI;
    const double factor_W_2D = inv_h_k * inv_h_k * 7.0 / (4.0 * Pi);
    const double factor_W_3D = inv_h_k * inv_h_k * inv_h_k * 21.0 / (16.0 * Pi);
    const double dimfac = D == 2 ? factor_W_2D * h * h : factor_W_3D * h * h * h;
    const double kernel_size = 2.0;
    const double dq = kernel_size / 20.0;
    double w[24], dw[24];
    for (int i = 0; i < 24; ++i)
    {
        double q = double(i - 1) * dq;
        w[i] = std::pow(1.0 - 0.5 * q, 4) * (1.0 + 2.0 * q);
        dw[i] = 0.625 * std::pow(q - 2.0, 3) * q;
    }
    // Neighbor<SPHAdaptation,SPHAdaptation> factors, neighbor_method.hpp:66-95
    const double inv_h = rel->inv_h, src_inv_h = rel->src_inv_h;
    const double inv_h2 = inv_h * inv_h, inv_h3 = inv_h2 * inv_h, inv_h4 = inv_h3 * inv_h;
    const double src_inv_h2 = src_inv_h * src_inv_h, src_inv_h3 = src_inv_h2 * src_inv_h;
    const double fW = (D == 2 ? inv_h2 : inv_h3) * dimfac;
    const double fdW = (D == 2 ? inv_h3 : inv_h4) * dimfac;
    KernelPoly<T> kp;
    expand_table(w, (long double)fW, kp.w, dq);
    expand_table(dw, (long double)fdW, kp.dw, dq);
    kp.inv_h = T(inv_h);
    kp.x_per_r = T(inv_h * (1.0 / dq));
    // W0 = src_inv_h^D dimfac W(0); interpolateCubic at q = 0 returns data[1] scaled by the basis
    const double f1 = 0.0, f0 = dq, f2 = -dq, f3 = -2 * dq;
    const double dl0 = (-1.0 * dq) * (-2.0 * dq) * (-3.0 * dq), dl1 = dq * (-1.0 * dq) * (-2.0 * dq);
    const double dl2 = (2.0 * dq) * dq * (-1.0 * dq), dl3 = (3.0 * dq) * (2.0 * dq) * dq;
    const double interp0 = (f1 * f2 * f3) / dl0 * w[0] + (f0 * f2 * f3) / dl1 * w[1] +
                           (f0 * f1 * f3) / dl2 * w[2] + (f0 * f1 * f2) / dl3 * w[3];
    kp.w0 = T((D == 2 ? src_inv_h2 : src_inv_h3) * dimfac * interp0);
    if (!rel->poly)
    {
        sphx_status s = dev_alloc(src->ctx->stream, (void **)&rel->poly, sizeof(KernelPoly<T>));
        if (s)
            return s;
    }
    SPHX_HIP(hipMemcpyAsync(rel->poly, &kp, sizeof(kp), hipMemcpyHostToDevice, src->ctx->stream));
    SPHX_HIP(hipStreamSynchronize(src->ctx->stream));
    return SPHX_OK;
}

sphx_status build_kernel_poly(sphx_relation *rel)
{
    return rel->source->P == P_FP64 ? build_poly_t<double>(rel) : build_poly_t<float>(rel);
}

sphx_status step_begin(sphx_context *ctx, double u_ref, double advection_cfl, double h, double time)
{
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->scalars_d, u_ref, advection_cfl, h,
                       time);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

sphx_status substep_begin(sphx_context *ctx, double acoustic_cfl, double h)
{
    hipLaunchKernelGGL(k_substep_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->scalars_d, acoustic_cfl, h);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

//=============================================================================
// batched launches (host side)
//=============================================================================
size_t batch_scene_args_bytes(int P) { return P == P_MIXED ? sizeof(SceneArgs<float>) : sizeof(SceneArgs<double>); }

template <class T>
static void fill_scene_args(void *dst, sphx_relation *inner, sphx_relation *const *w, int nw,
                            const sphx_advection_config *cfg, const double *time_dev, bool staged)
{
    sphx_body *b = inner->source;
    SceneArgs<T> a{};
    a.f = fluid_args<T>(b);
    a.f.step = b->ctx->scalars_d + SLOT_STEP;
    a.f.lanes = b->ctx->scalars_d + SLOT_MAX;
    a.f.riemann = cfg->riemann;
    attach_relations(a.f, inner, w, nw);
    if (!staged)
    {
        a.f.stage_count = nullptr;
        a.f.stage_keys = nullptr;
        a.f.local_slice = nullptr;
        a.f.stage_cap = 0;
    }
    a.pa = pos_args(b);
    a.Vol = (T *)b->Vol;
    a.posvol = (Rec4<T> *)b->posvol;
    a.pos_f32 = b->pos_f32;
    a.scalars = b->ctx->scalars_d;
    a.time = time_dev;
    a.u_ref = cfg->u_ref;
    a.advection_cfl = cfg->advection_cfl;
    a.acoustic_cfl = cfg->acoustic_cfl;
    a.h = b->h;
    a.flow = cfg->flow_type;
    std::memcpy(dst, &a, sizeof(a));
}

void batch_scene_args(void *dst, sphx_relation *inner, sphx_relation *const *w, int nw,
                      const sphx_advection_config *cfg, const double *time_dev, bool staged)
{
    if (inner->source->P == P_MIXED)
        fill_scene_args<float>(dst, inner, w, nw, cfg, time_dev, staged);
    else
        fill_scene_args<double>(dst, inner, w, nw, cfg, time_dev, staged);
}

template <int D, int P, bool RIEMANN>
struct BatchL
{
    using T = typename Prec<P>::T;
    static const SceneArgs<T> *tab(const BatchFluid &b) { return (const SceneArgs<T> *)b.tab; }
    // dynamic LDS of the staged neighbour passes (as the single-scene launchers)
    static size_t lds_sum(const BatchFluid &b) { return (size_t)b.stage_cap * sizeof(Rec4<T>); }
    static size_t lds_pass1(const BatchFluid &b)
    {
        return b.stage_cap ? (size_t)pass1_cap<P>() * (sizeof(Rec4<T>) + sizeof(T)) : 0;
    }
    static size_t lds_pass2(const BatchFluid &b)
    {
        return b.stage_cap ? (size_t)kPass2Cap * (sizeof(Rec4<T>) + 3 * sizeof(T)) : 0;
    }
    static sphx_status head(const BatchFluid &b)
    {
        const dim3 g(b.blocks), blk(kThreads), gs(grid_for(b.n_scenes, 64)), bs(64);
        allow_lds(kb_summation<D, P>, lds_sum(b));
        hipLaunchKernelGGL((kb_summation<D, P>), g, blk, lds_sum(b), b.stream, tab(b), b.map);
        hipLaunchKernelGGL((kb_regularize_setup<D, P>), g, blk, 0, b.stream, tab(b), b.map);
        hipLaunchKernelGGL((kb_reset_lanes<T>), dim3(b.n_scenes), dim3(64), 0, b.stream, tab(b), b.n_scenes);
        hipLaunchKernelGGL((kb_speed_max<D, P>), g, blk, 0, b.stream, tab(b), b.map);
        hipLaunchKernelGGL((kb_step_begin<T>), gs, bs, 0, b.stream, tab(b), b.n_scenes);
        hipLaunchKernelGGL((kb_signal_max<D, P>), g, blk, 0, b.stream, tab(b), b.map);
        SPHX_HIP(hipGetLastError());
        return SPHX_OK;
    }
    static sphx_status substep(const BatchFluid &b)
    {
        const dim3 g(b.blocks), blk(kThreads), gs(grid_for(b.n_scenes, 64)), bs(64);
        allow_lds(kb_pass1<D, P, RIEMANN>, lds_pass1(b));
        allow_lds(kb_pass2<D, P, RIEMANN>, lds_pass2(b));
        hipLaunchKernelGGL((kb_substep_begin<T>), gs, bs, 0, b.stream, tab(b), b.n_scenes);
        hipLaunchKernelGGL((kb_init_1st<D, P>), g, blk, 0, b.stream, tab(b), b.map);
        hipLaunchKernelGGL((kb_pass1<D, P, RIEMANN>), g, blk, lds_pass1(b), b.stream, tab(b), b.map);
        hipLaunchKernelGGL((kb_pass2<D, P, RIEMANN>), g, blk, lds_pass2(b), b.stream, tab(b), b.map);
        SPHX_HIP(hipGetLastError());
        return SPHX_OK;
    }
    static sphx_status gate(const BatchFluid &b)
    {
        hipLaunchKernelGGL((kb_gate<T>), dim3(grid_for(b.n_scenes, 64)), dim3(64), 0, b.stream, tab(b), b.n_scenes,
                           b.tail_on, b.gate, b.state);
        SPHX_HIP(hipGetLastError());
        return SPHX_OK;
    }
    static sphx_status position(const BatchFluid &b)
    {
        hipLaunchKernelGGL((kb_update_position<D, P>), dim3(b.blocks), dim3(kThreads), 0, b.stream, tab(b), b.map,
                           b.gate);
        SPHX_HIP(hipGetLastError());
        return SPHX_OK;
    }
};

template <template <int, int, bool> class L>
struct BatchDispatch
{
    template <class F>
    static sphx_status run(const BatchFluid &b, F &&f)
    {
        const bool rs = b.riemann != SPHX_RIEMANN_NONE;
#define SPHX_BATCH_CASE(DD, PP)                                                                                   \
    if (b.D == DD && b.P == PP)                                                                                  \
        return rs ? f(L<DD, PP, true>{}) : f(L<DD, PP, false>{});
        SPHX_BATCH_CASE(2, P_FP64)
        SPHX_BATCH_CASE(3, P_FP64)
        SPHX_BATCH_CASE(2, P_MIXED)
        SPHX_BATCH_CASE(3, P_MIXED)
#undef SPHX_BATCH_CASE
        set_error("batched launch: unsupported dimension or precision");
        return SPHX_E_ARG;
    }
};

sphx_status batch_step_head(const BatchFluid &b)
{
    return BatchDispatch<BatchL>::run(b, [&](auto l) { return decltype(l)::head(b); });
}
sphx_status batch_substep(const BatchFluid &b)
{
    return BatchDispatch<BatchL>::run(b, [&](auto l) { return decltype(l)::substep(b); });
}
sphx_status batch_step_gate(const BatchFluid &b)
{
    return BatchDispatch<BatchL>::run(b, [&](auto l) { return decltype(l)::gate(b); });
}
sphx_status batch_position(const BatchFluid &b)
{
    return BatchDispatch<BatchL>::run(b, [&](auto l) { return decltype(l)::position(b); });
}

} // namespace sphx
