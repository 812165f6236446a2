/*
 * sphx_oracle.hpp -- CPU restatement of the SPHinXsys computing-kernel (CK)
 * weakly-compressible SPH step.  TEST INFRASTRUCTURE ONLY: this code is the
 * parity checker and the CPU baseline.  The product library never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line whose arithmetic it restates
 * (paths relative to the SPHinXsys source tree).  Expression order follows
 * the reference so that FP64 results differ from it only by the summation
 * order of neighbour sums and by Eigen's reduction order for dot/squaredNorm
 * (taken here as ((x*x + y*y) + z*z)); both are below the 1e-10 relative gate.
 *
 * Canonical orders (the reference fills these with relaxed atomics, so its
 * order is not reproducible; sets are identical):
 *  - cell list: ascending particle index inside each cell;
 *  - inner/contact neighbour list of i: cells in mesh_for_each order
 *    (l -> m -> n, src/for_3D_build/meshes/mesh_iterators.hpp:19-27), then
 *    cell-list order inside each cell;
 *  - particle sort: stable sort of Morton keys (the reference's quicksort is
 *    unstable, so only the key sequence is comparable to it).
 */
#ifndef SPHX_ORACLE_HPP
#define SPHX_ORACLE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

namespace orc
{
using UInt = uint32_t;

//---------------------------------------------------------------------------
// threading: static chunks over [0,n), standing in for tbb::parallel_for
//---------------------------------------------------------------------------
int &thread_count();

template <class F>
void parallel_for(size_t n, const F &f)
{
    int nt = thread_count();
    if (nt <= 1 || n < 4096)
    {
        for (size_t i = 0; i < n; ++i)
            f(i);
        return;
    }
    std::vector<std::thread> pool;
    size_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t)
    {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e)
            break;
        pool.emplace_back([=, &f]() { for (size_t i = b; i < e; ++i) f(i); });
    }
    for (auto &th : pool)
        th.join();
}

//---------------------------------------------------------------------------
// small fixed vectors (Eigen Vecd stand-in)
//---------------------------------------------------------------------------
template <class R, int D>
struct Vec
{
    R v[D];
    R &operator[](int k) { return v[k]; }
    const R &operator[](int k) const { return v[k]; }
    static Vec Zero()
    {
        Vec r;
        for (int k = 0; k < D; ++k)
            r.v[k] = R(0);
        return r;
    }
};
template <class R, int D>
inline Vec<R, D> operator-(const Vec<R, D> &a, const Vec<R, D> &b)
{
    Vec<R, D> r;
    for (int k = 0; k < D; ++k)
        r[k] = a[k] - b[k];
    return r;
}
template <class R, int D>
inline Vec<R, D> operator+(const Vec<R, D> &a, const Vec<R, D> &b)
{
    Vec<R, D> r;
    for (int k = 0; k < D; ++k)
        r[k] = a[k] + b[k];
    return r;
}
template <class R, int D>
inline Vec<R, D> operator*(const Vec<R, D> &a, R s)
{
    Vec<R, D> r;
    for (int k = 0; k < D; ++k)
        r[k] = a[k] * s;
    return r;
}
template <class R, int D>
inline Vec<R, D> operator*(R s, const Vec<R, D> &a)
{
    Vec<R, D> r;
    for (int k = 0; k < D; ++k)
        r[k] = s * a[k];
    return r;
}
template <class R, int D>
inline Vec<R, D> operator/(const Vec<R, D> &a, R s)
{
    Vec<R, D> r;
    for (int k = 0; k < D; ++k)
        r[k] = a[k] / s;
    return r;
}
template <class R, int D>
inline Vec<R, D> &operator+=(Vec<R, D> &a, const Vec<R, D> &b)
{
    for (int k = 0; k < D; ++k)
        a[k] += b[k];
    return a;
}
template <class R, int D>
inline Vec<R, D> &operator-=(Vec<R, D> &a, const Vec<R, D> &b)
{
    for (int k = 0; k < D; ++k)
        a[k] -= b[k];
    return a;
}
// Eigen redux order for dot/squaredNorm is taken as ((a0 + a1) + a2).
template <class R, int D>
inline R dot(const Vec<R, D> &a, const Vec<R, D> &b)
{
    R s = a[0] * b[0];
    for (int k = 1; k < D; ++k)
        s += a[k] * b[k];
    return s;
}
template <class R, int D>
inline R squaredNorm(const Vec<R, D> &a) { return dot(a, a); }
template <class R, int D>
inline R norm(const Vec<R, D> &a) { return std::sqrt(squaredNorm(a)); }
// Eigen MatrixBase::normalized(): n / sqrt(|n|^2) if |n|^2 > 0, else n.
template <class R, int D>
inline Vec<R, D> normalized(const Vec<R, D> &a)
{
    R z = squaredNorm(a);
    return z > R(0) ? a / R(std::sqrt(z)) : a;
}
//---------------------------------------------------------------------------
// small fixed matrices (Eigen Matd stand-in), row-major m[row][col]
//---------------------------------------------------------------------------
template <class R, int D>
struct Mat
{
    R m[D][D];
    static Mat Zero()
    {
        Mat r;
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b)
                r.m[a][b] = R(0);
        return r;
    }
    static Mat Identity()
    {
        Mat r = Zero();
        for (int a = 0; a < D; ++a)
            r.m[a][a] = R(1);
        return r;
    }
    Mat transpose() const
    {
        Mat r;
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b)
                r.m[a][b] = m[b][a];
        return r;
    }
};
template <class R, int D>
inline Mat<R, D> operator+(const Mat<R, D> &x, const Mat<R, D> &y)
{
    Mat<R, D> r;
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
            r.m[a][b] = x.m[a][b] + y.m[a][b];
    return r;
}
template <class R, int D>
inline Mat<R, D> operator*(const Mat<R, D> &x, R s)
{
    Mat<R, D> r;
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
            r.m[a][b] = x.m[a][b] * s;
    return r;
}
template <class R, int D>
inline Mat<R, D> operator*(R s, const Mat<R, D> &x)
{
    Mat<R, D> r;
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
            r.m[a][b] = s * x.m[a][b];
    return r;
}
// matrix product; the inner sum runs in index order
template <class R, int D>
inline Mat<R, D> operator*(const Mat<R, D> &x, const Mat<R, D> &y)
{
    Mat<R, D> r;
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
        {
            R s = x.m[a][0] * y.m[0][b];
            for (int k = 1; k < D; ++k)
                s += x.m[a][k] * y.m[k][b];
            r.m[a][b] = s;
        }
    return r;
}
template <class R, int D>
inline Vec<R, D> operator*(const Mat<R, D> &x, const Vec<R, D> &v)
{
    Vec<R, D> r;
    for (int a = 0; a < D; ++a)
    {
        R s = x.m[a][0] * v[0];
        for (int k = 1; k < D; ++k)
            s += x.m[a][k] * v[k];
        r[a] = s;
    }
    return r;
}
template <class R, int D>
inline Mat<R, D> &operator-=(Mat<R, D> &x, const Mat<R, D> &y)
{
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
            x.m[a][b] -= y.m[a][b];
    return x;
}
// outer product u v^T
template <class R, int D>
inline Mat<R, D> outer(const Vec<R, D> &u, const Vec<R, D> &v)
{
    Mat<R, D> r;
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b)
            r.m[a][b] = u[a] * v[b];
    return r;
}
// determinant by cofactor expansion along the first column
template <class R, int D>
inline R determinant(const Mat<R, D> &x)
{
    if (D == 2)
        return x.m[0][0] * x.m[1][1] - x.m[1][0] * x.m[0][1];
    auto minor = [&](int a, int b, int c) {
        return x.m[a][0] * (x.m[b][1] * x.m[c][2] - x.m[b][2] * x.m[c][1]);
    };
    return minor(0, 1, 2) - minor(1, 0, 2) + minor(2, 0, 1);
}
// inverse as adjugate / determinant (the determinant from the first-column cofactors)
template <class R, int D>
inline Mat<R, D> inverse(const Mat<R, D> &x)
{
    Mat<R, D> r;
    if (D == 2)
    {
        const R inv_det = R(1) / (x.m[0][0] * x.m[1][1] - x.m[1][0] * x.m[0][1]);
        r.m[0][0] = x.m[1][1] * inv_det;
        r.m[1][0] = -x.m[1][0] * inv_det;
        r.m[0][1] = -x.m[0][1] * inv_det;
        r.m[1][1] = x.m[0][0] * inv_det;
        return r;
    }
    // cofactor of entry (i, j) with cyclic indices
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return x.m[i1][j1] * x.m[i2][j2] - x.m[i1][j2] * x.m[i2][j1];
    };
    const R det = cof(0, 0) * x.m[0][0] + cof(1, 0) * x.m[1][0] + cof(2, 0) * x.m[2][0];
    const R inv_det = R(1) / det;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            r.m[a][b] = cof(b, a) * inv_det;
    return r;
}
// Tikhonov-regularised inverse (B^T B + eps I)^-1 B^T, common/vector_functions.h:198-203
template <class R, int D>
inline Mat<R, D> inverseTikhonov(const Mat<R, D> &x, R eps)
{
    const Mat<R, D> xt = x.transpose();
    return inverse(xt * x + eps * Mat<R, D>::Identity()) * xt;
}

// src/shared/common/scalar_functions.h:128-131
template <class R>
inline R SGN(R x) { return (x < 0) ? R(-1) : ((x > 0) ? R(1) : R(0)); }

template <class R>
constexpr R TinyReal() { return R(2.71051e-20); } // base_data_type.h:207
template <class R>
constexpr R SqrtEps() { return R(1.0e-8); } // base_data_type.h:206

//---------------------------------------------------------------------------
// Wendland C2 kernel and its 24-entry tabulation
// reference: kernels/kernel_wendland_c2.cpp:8-60, kernels/base_kernel.cpp:6-27,
//            kernels/base_kernel.h:91-93,
//            shared_ck/smoothing_kernel/kernel_tabulated_ck.{h:37-58,cpp:6-25}
//---------------------------------------------------------------------------
constexpr int kernel_resolution = 20;
constexpr int tabulated_size = kernel_resolution + 4;

template <class R>
struct TabulatedKernel
{
    R h, inv_h, kernel_size;
    R factor_W_1D, factor_W_2D, factor_W_3D;
    R dimension_factor_1D, dimension_factor_2D, dimension_factor_3D;
    R dq, delta_q_0, delta_q_1, delta_q_2, delta_q_3;
    R w[tabulated_size], dw[tabulated_size], d2w[tabulated_size];

    static R W_1D(R q) { return std::pow(1.0 - 0.5 * q, 4) * (1.0 + 2.0 * q); }
    static R dW_1D(R q) { return 0.625 * std::pow(q - 2.0, 3) * q; }
    static R d2W_1D(R q) { return 1.25 * std::pow(q - 2.0, 2) * (2.0 * q - 1.0); }

    explicit TabulatedKernel(R h_in)
    {
        const R Pi = R(M_PI);
        h = h_in;
        inv_h = 1.0 / h;
        kernel_size = 2.0;
        factor_W_1D = inv_h * 3.0 / 4.0;
        factor_W_2D = inv_h * inv_h * 7.0 / (4.0 * Pi);
        factor_W_3D = inv_h * inv_h * inv_h * 21.0 / (16.0 * Pi);
        dimension_factor_1D = factor_W_1D * h;
        dimension_factor_2D = factor_W_2D * h * h;
        dimension_factor_3D = factor_W_3D * h * h * h;
        dq = kernel_size / R(kernel_resolution);
        for (int i = 0; i < tabulated_size; i++)
        {
            w[i] = W_1D(R(i - 1) * dq);
            dw[i] = dW_1D(R(i - 1) * dq);
            d2w[i] = d2W_1D(R(i - 1) * dq);
        }
        delta_q_0 = (-1.0 * dq) * (-2.0 * dq) * (-3.0 * dq);
        delta_q_1 = dq * (-1.0 * dq) * (-2.0 * dq);
        delta_q_2 = (2.0 * dq) * dq * (-1.0 * dq);
        delta_q_3 = (3.0 * dq) * (2.0 * dq) * dq;
    }
    // kernel_tabulated_ck.h:45-58
    R interpolateCubic(const R *data, R q) const
    {
        int location = (int)std::floor(q / dq);
        int i = location + 1;
        R fraction_1 = q - R(location) * dq;
        R fraction_0 = fraction_1 + dq;
        R fraction_2 = fraction_1 - dq;
        R fraction_3 = fraction_1 - 2 * dq;
        return (fraction_1 * fraction_2 * fraction_3) / delta_q_0 * data[i - 1] +
               (fraction_0 * fraction_2 * fraction_3) / delta_q_1 * data[i] +
               (fraction_0 * fraction_1 * fraction_3) / delta_q_2 * data[i + 1] +
               (fraction_0 * fraction_1 * fraction_2) / delta_q_3 * data[i + 2];
    }
    R dimension_factor(int d) const { return d == 2 ? dimension_factor_2D : dimension_factor_3D; }
};

//---------------------------------------------------------------------------
// Neighbor<SPHAdaptation,SPHAdaptation> smoothing kernel and criterion
// reference: shared_ck/body_relation/neighbor_method.hpp:18-161
//---------------------------------------------------------------------------
template <class R, int D>
struct NeighborKernel
{
    const TabulatedKernel<R> *k;
    R src_inv_h, src_inv_h_squared, src_inv_h_cubed;
    R inv_h, inv_h_squared, inv_h_cubed, inv_h_fourth, inv_h_fifth;
    R kernel_size_squared;

    NeighborKernel(const TabulatedKernel<R> &kern, R src_h, R tar_h) : k(&kern)
    {
        src_inv_h = 1.0 / src_h;
        inv_h = 1.0 / std::max(src_h, tar_h);
        src_inv_h_squared = src_inv_h * src_inv_h;
        src_inv_h_cubed = src_inv_h_squared * src_inv_h;
        inv_h_squared = inv_h * inv_h;
        inv_h_cubed = inv_h_squared * inv_h;
        inv_h_fourth = inv_h_cubed * inv_h;
        inv_h_fifth = inv_h_fourth * inv_h;
        kernel_size_squared = std::pow(kern.kernel_size, R(2));
    }
    R dimfac() const { return k->dimension_factor(D); }
    // neighbor_method.hpp:111-122 (W), 124-135 (dW), 97-108 (W0); 18-58 (factors)
    R W(const Vec<R, D> &disp) const
    {
        R scale = D == 2 ? inv_h_squared : inv_h_cubed;
        return scale * dimfac() * k->interpolateCubic(k->w, norm(disp) * inv_h);
    }
    R dW(const Vec<R, D> &disp) const
    {
        R scale = D == 2 ? inv_h_cubed : inv_h_fourth;
        return scale * dimfac() * k->interpolateCubic(k->dw, norm(disp) * inv_h);
    }
    R W0() const
    {
        R scale = D == 2 ? src_inv_h_squared : src_inv_h_cubed;
        return scale * dimfac() * k->interpolateCubic(k->w, R(0));
    }
    // neighbor_method.hpp:152-156; note the (j, i) argument order of the reference
    bool criterion(const Vec<R, D> &src_pos_i, const Vec<R, D> &tar_pos_j) const
    {
        return squaredNorm(inv_h * (src_pos_i - tar_pos_j)) < kernel_size_squared;
    }
};

//---------------------------------------------------------------------------
// Mesh: reference meshes/base_mesh.cpp:6-16, base_mesh.hxx:9-26,73-99
//---------------------------------------------------------------------------
template <class R, int D>
struct Mesh
{
    Vec<R, D> lower;
    R gs;
    int all_grid_points[D];
    int all_cells[D];
    UInt linear_offset = 0;

    Mesh() = default;
    Mesh(const Vec<R, D> &bl, const Vec<R, D> &bu, R grid_spacing, UInt buffer_width)
    {
        gs = grid_spacing;
        R buffer = R(buffer_width) * grid_spacing;
        for (int k = 0; k < D; ++k)
        {
            lower[k] = bl[k] - buffer;
            R tentative_dimension = bu[k] + buffer - lower[k];
            all_grid_points[k] = (int)std::ceil(tentative_dimension / grid_spacing) + 1;
            all_cells[k] = all_grid_points[k] - 1;
        }
    }
    size_t NumberOfCells() const
    {
        size_t n = 1;
        for (int k = 0; k < D; ++k)
            n *= (size_t)all_cells[k];
        return n;
    }
    void CellIndexFromPosition(const Vec<R, D> &p, int *c) const
    {
        for (int k = 0; k < D; ++k)
        {
            int ci = (int)std::floor((p[k] - lower[k]) / gs);
            c[k] = std::min(std::max(ci, 0), all_grid_points[k] - 2);
        }
    }
    UInt LinearCellIndex(const int *c) const
    {
        if (D == 2)
            return linear_offset + (UInt)(c[0] * all_cells[1] + c[1]);
        return linear_offset + (UInt)(c[0] * all_cells[1] * all_cells[2] + c[1] * all_cells[2] + c[2]);
    }
    UInt LinearCellIndexFromPosition(const Vec<R, D> &p) const
    {
        int c[D];
        CellIndexFromPosition(p, c);
        return LinearCellIndex(c);
    }
    // base_mesh.hxx:91-99
    static UInt MortonCode(UInt i)
    {
        UInt x = i;
        x &= 0x3ff;
        x = (x | x << 16) & 0x30000ff;
        x = (x | x << 8) & 0x300f00f;
        x = (x | x << 4) & 0x30c30c3;
        x = (x | x << 2) & 0x9249249;
        return x;
    }
    static UInt MortonOrder(const int *c)
    {
        if (D == 2)
            return MortonCode(c[0]) | (MortonCode(c[1]) << 1);
        return MortonCode(c[0]) | (MortonCode(c[1]) << 1) | (MortonCode(c[2]) << 2);
    }
    // CellPositionFromIndex (base_mesh.cpp:23-26, base_mesh.hxx:28-31)
    Vec<R, D> CellPosition(const int *c) const
    {
        Vec<R, D> r;
        for (int k = 0; k < D; ++k)
            r[k] = lower[k] + R(c[k]) * gs + R(0.5) * gs;
        return r;
    }
};

// mesh_for_each over [lo, hi) with the last axis fastest
template <int D, class F>
inline void mesh_for_each(const int *lo, const int *hi, const F &f)
{
    if (D == 2)
    {
        for (int l = lo[0]; l < hi[0]; ++l)
            for (int m = lo[1]; m < hi[1]; ++m)
            {
                int c[2] = {l, m};
                f(c);
            }
    }
    else
    {
        for (int l = lo[0]; l < hi[0]; ++l)
            for (int m = lo[1]; m < hi[1]; ++m)
                for (int n = lo[2]; n < hi[2]; ++n)
                {
                    int c[3] = {l, m, n};
                    f(c);
                }
    }
}

//---------------------------------------------------------------------------
// exclusive scan returning the total (algorithm_primitive.h:244-279)
//---------------------------------------------------------------------------
inline UInt exclusive_scan(const UInt *in, UInt *out, size_t n)
{
    UInt acc = 0;
    for (size_t i = 0; i < n; ++i)
    {
        UInt v = in[i];
        out[i] = acc;
        acc += v;
    }
    return acc;
}

//---------------------------------------------------------------------------
// Riemann solver constants, riemann_solver_ck.hpp:59-69 (single fluid: i == j)
//---------------------------------------------------------------------------
template <class R>
struct Impedance
{
    R rho0, rho0c0, inv_rho0c0_sum, inv_rho0c0_ave, rho0c0_geo_ave, inv_c0_ave;
    Impedance(R rho0_in, R c0)
    {
        rho0 = rho0_in;
        rho0c0 = rho0 * c0;
        inv_rho0c0_sum = 1.0 / (rho0c0 + rho0c0);
        inv_rho0c0_ave = (rho0c0 + rho0c0) / (std::pow(rho0c0, R(2)) + std::pow(rho0c0, R(2)));
        rho0c0_geo_ave = 2.0 * rho0c0 * rho0c0 * inv_rho0c0_sum;
        inv_c0_ave = 0.5 * (rho0 + rho0) * inv_rho0c0_ave;
    }
};

} // namespace orc
#endif
