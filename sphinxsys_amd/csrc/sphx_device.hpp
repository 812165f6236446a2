// sphx_device.hpp -- device-side types and helpers shared by the gfx950 kernels.
//
// Data layout in HBM (per body, structure of arrays):
//   pos_x/pos_y/pos_z, dpos_*      double            (both precision modes)
//   posvol[i]                      Rec4<T>  = (x, y, z|vol, vol|pad) gather record
//   vel[i]                         Rec4<T>  = (vx, vy, vz, 0) gather record
//   force_*, force_prior_*         T
//   Vol, mass, rho, p, C, dC, ...  T
// with T = double (FP64 mode) or float (MIXED mode).  In MIXED mode the
// compression field stores C - 1 so that float keeps the ~1e-3 density
// variation of a weakly-compressible fluid to full relative precision.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphx
{
enum
{
    P_FP64 = 0,
    P_MIXED = 1
};

template <class T>
struct Rec4;
template <>
struct alignas(16) Rec4<float>
{
    float x, y, z, w;
};
template <>
struct alignas(32) Rec4<double>
{
    double x, y, z, w;
};

template <int P>
struct Prec;
template <>
struct Prec<P_FP64>
{
    using T = double;
    static constexpr bool c_minus_one = false;
};
template <>
struct Prec<P_MIXED>
{
    using T = float;
    static constexpr bool c_minus_one = true;
};

// Tabulated Wendland C2 kernel re-expressed per interval as a cubic in the
// fraction f = x - loc of the table coordinate x = q / dq = r inv_h / dq (the
// Lagrange form of kernel_tabulated_ck.h:45-58 expanded on the host),
// premultiplied by the dimension and smoothing-length factors.
constexpr int kPolyIntervals = 21; // loc in [0, 20]; q in [0, 2.1)

template <class T>
struct KernelPoly
{
    T w[kPolyIntervals][4];  // W  = sum_k w[loc][k]  f^k   (factor inv_h^D dimfac folded in)
    T dw[kPolyIntervals][4]; // dW = sum_k dw[loc][k] f^k   (factor inv_h^(D+1) dimfac folded in)
    T inv_h, x_per_r, w0;    // x_per_r = inv_h / dq; w0 = W0 of the source with its own factor
};

// Mesh of a cell linked list, base_mesh.cpp:6-16 / base_mesh.hxx:9-26.
struct MeshParams
{
    double lower[3];
    double gs;
    int all_grid_points[3];
    int all_cells[3];
};

// Mesh::CellIndexFromPosition and the linear cell index, base_mesh.hxx:9-26,73-78
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

// the double positions of a body (SoA)
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

// XCD-aware block index: gfx950 dispatches consecutive workgroups round-robin
// over its 8 XCDs; remap so that each XCD works on one contiguous range of
// particles and the neighbour gathers of a range stay in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block_index()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t per = nb / 8u, rem = nb % 8u;
    const uint32_t xcd = b % 8u, slot = b / 8u;
    // blocks [0, rem*(per+1)) belong to the first `rem` XCDs with per+1 blocks each
    const uint32_t base = xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per;
    return base + slot;
}

template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }
template <class T>
__device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }

// order-preserving maps of floating point values onto unsigned integers so that
// a global max can be taken with integer atomics
__device__ __forceinline__ unsigned long long ordered_bits(double v)
{
    unsigned long long u = __double_as_longlong(v);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double from_ordered_bits(unsigned long long u)
{
    unsigned long long r = (u & 0x8000000000000000ull) ? (u & ~0x8000000000000000ull) : ~u;
    double d;
    __builtin_memcpy(&d, &r, sizeof(d));
    return d;
}

// Max reductions accumulate with atomicMax on the ordered bits of a double.
// Device-scope atomics on one address serialise in memory, so each workgroup
// uses one of kMaxLanes accumulators, kMaxStride u64 (one 128-B line) apart,
// and the reader takes the maximum of the lanes.
constexpr unsigned long long kOrderedLowest = 0x0010000000000000ull; // ordered_bits(-DBL_MAX)
constexpr int kMaxLanes = 32;
constexpr int kMaxStride = 16;
__device__ __forceinline__ void atomic_max_spread(unsigned long long *base, double v)
{
    atomicMax(base + (blockIdx.x % kMaxLanes) * kMaxStride, ordered_bits(v));
}

// wave64 reductions with DPP-free shuffles
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

} // namespace sphx
