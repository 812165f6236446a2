/*
 * sphx_oracle.cpp -- the CPU restatement behind the sphx C ABI.
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); the product
 * library never links or loads it.  Precision SPHX_FP64 restates the
 * reference's Real=double build, precision 2 (ORACLE_FP32) its
 * SPHINXSYS_USE_FLOAT build; SPHX_MIXED is a device-only mode and is
 * rejected here.
 *
 * Reference citations are file:line relative to the SPHinXsys source tree.
 */
#include "../include/sphx.h"
#include "sphx_oracle.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

namespace orc
{
int &thread_count()
{
    static int n = []() {
        const char *e = std::getenv("SPHX_ORACLE_THREADS");
        if (e && std::atoi(e) > 0)
            return std::atoi(e);
        unsigned hc = std::thread::hardware_concurrency();
        return hc > 0 ? (int)hc : 1;
    }();
    return n;
}
} // namespace orc

using namespace orc;

static thread_local std::string g_last_error;
static sphx_status fail(sphx_status s, const std::string &msg)
{
    g_last_error = msg;
    return s;
}

enum { ORACLE_FP32 = 2 };

//=============================================================================
// objects
//=============================================================================
struct BodyBase
{
    sphx_body_desc desc;
    bool vol_written = false;
    virtual ~BodyBase() = default;
};

template <class R, int D>
struct Body : BodyBase
{
    using Real = R;
    static constexpr int Dim = D;
    using V = Vec<R, D>;
    UInt n;
    R h;
    std::unique_ptr<TabulatedKernel<R>> kernel;
    std::vector<V> pos, vel, force, force_prior, dpos, normal, wall_vel, wall_acc;
    std::vector<R> Vol, mass, rho, p, C, dC, Vol_ref, S, interp, pos_div;
    std::vector<UInt> original_id, sorted_id, sequence;
    std::vector<Mat<R, D>> correction;         // "LinearCorrectionMatrix", initial identity
    std::vector<int32_t> indicator, prev_indicator; // "Indicator" (0), "PreviousSurfaceIndicator" (1)
    std::vector<V> viscous_force, prev_viscous_force, kernel_gradient_integral;

    explicit Body(const sphx_body_desc &d)
    {
        desc = d;
        n = d.n_particles;
        // adaptation.cpp:12-19: spacing_ref = resolution / refinement(1), h = ratio * spacing_ref
        R spacing_ref = R(d.spacing_ref) / R(1.0);
        h = R(d.h_spacing_ratio) * spacing_ref;
        kernel.reset(new TabulatedKernel<R>(h));
        V z = V::Zero();
        pos.assign(n, z), vel.assign(n, z), force.assign(n, z), force_prior.assign(n, z);
        dpos.assign(n, z), normal.assign(n, z), wall_vel.assign(n, z), wall_acc.assign(n, z);
        Vol.assign(n, R(0)), mass.assign(n, R(0)), rho.assign(n, R(d.rho0)), p.assign(n, R(0));
        C.assign(n, R(1)), dC.assign(n, R(0)), Vol_ref.assign(n, R(0)), S.assign(n, R(1));
        interp.assign(n, R(0)), pos_div.assign(n, R(0));
        original_id.resize(n), sorted_id.resize(n), sequence.assign(n, 0);
        // kernel_correction_ck.hpp:12-15; surface_indication_ck.hpp:17-18, 41-44
        correction.assign(n, Mat<R, D>::Identity());
        indicator.assign(n, 0), prev_indicator.assign(n, 1);
        viscous_force.assign(n, z), prev_viscous_force.assign(n, z), kernel_gradient_integral.assign(n, z);
        for (UInt i = 0; i < n; ++i)
            original_id[i] = sorted_id[i] = i;
    }
};

struct sphx_context
{
    int device;
};
struct sphx_body
{
    sphx_context *ctx;
    std::unique_ptr<BodyBase> impl;
};

struct CLLBase
{
    virtual ~CLLBase() = default;
};
template <class R, int D>
struct CLL : CLLBase
{
    Mesh<R, D> mesh;
    std::vector<UInt> cell_offset, particle_index;
    bool built = false;
};
struct sphx_cell_linked_list
{
    sphx_body *body;
    std::unique_ptr<CLLBase> impl;
};

struct sphx_relation
{
    bool inner;
    sphx_body *source, *target;
    sphx_cell_linked_list *cll; // of the target
    std::vector<UInt> neighbor_size, particle_offset, neighbor_index;
    bool built = false;
};

#define BODY_TYPES(B)                                        \
    using R = typename std::decay<decltype(B)>::type::Real; \
    constexpr int D = std::decay<decltype(B)>::type::Dim

//=============================================================================
// dispatch on (precision, dim)
//=============================================================================
template <class F>
static sphx_status with_body(sphx_body *b, F &&f)
{
    if (!b)
        return fail(SPHX_E_ARG, "null body");
    int d = b->impl->desc.dim, pr = b->impl->desc.precision;
    if (pr == SPHX_FP64 && d == 2)
        return f(*static_cast<Body<double, 2> *>(b->impl.get()));
    if (pr == SPHX_FP64 && d == 3)
        return f(*static_cast<Body<double, 3> *>(b->impl.get()));
    if (pr == ORACLE_FP32 && d == 2)
        return f(*static_cast<Body<float, 2> *>(b->impl.get()));
    if (pr == ORACLE_FP32 && d == 3)
        return f(*static_cast<Body<float, 3> *>(b->impl.get()));
    return fail(SPHX_E_ARG, "unsupported precision/dimension");
}
template <class R, int D>
static Body<R, D> &as(sphx_body *b) { return *static_cast<Body<R, D> *>(b->impl.get()); }
template <class R, int D>
static CLL<R, D> &as(sphx_cell_linked_list *c) { return *static_cast<CLL<R, D> *>(c->impl.get()); }

static bool same_kind(sphx_body *a, sphx_body *b)
{
    return a->impl->desc.dim == b->impl->desc.dim && a->impl->desc.precision == b->impl->desc.precision;
}

//=============================================================================
// algorithms
//=============================================================================
// UpdateCellLinkedList<P,RealBody>::exec, update_cell_linked_list.hpp:40-106.
// Count, exclusive scan, then fill in ascending particle index.
template <class R, int D>
static void update_cll(Body<R, D> &b, CLL<R, D> &c)
{
    size_t ncells = c.mesh.NumberOfCells();
    std::vector<UInt> count(ncells + 1, 0), cell_of(b.n);
    parallel_for(b.n, [&](size_t i) { cell_of[i] = c.mesh.LinearCellIndexFromPosition(b.pos[i]); });
    for (UInt i = 0; i < b.n; ++i)
        count[cell_of[i]]++;
    c.cell_offset.assign(ncells + 1, 0);
    exclusive_scan(count.data(), c.cell_offset.data(), ncells + 1);
    c.particle_index.assign(b.n, 0);
    std::vector<UInt> cursor(c.cell_offset.begin(), c.cell_offset.end() - 1);
    for (UInt i = 0; i < b.n; ++i)
        c.particle_index[cursor[cell_of[i]]++] = i;
    c.built = true;
}

// cells visited by NeighborSearch::forInnerSearch / forContactSearch with the
// one-cell search box, cell_linked_list.hpp:113-167
template <class R, int D, class F>
static void for_each_candidate(const CLL<R, D> &c, const Vec<R, D> &x, const F &f)
{
    int ci[D], lo[D], hi[D];
    c.mesh.CellIndexFromPosition(x, ci);
    for (int k = 0; k < D; ++k)
    {
        lo[k] = std::max(0, ci[k] - 1);
        hi[k] = std::min(c.mesh.all_cells[k], ci[k] + 2);
    }
    mesh_for_each<D>(lo, hi, [&](const int *cell) {
        UInt lin = c.mesh.LinearCellIndex(cell);
        for (UInt n = c.cell_offset[lin]; n < c.cell_offset[lin + 1]; ++n)
            f(c.particle_index[n]);
    });
}

// UpdateRelation<P,Inner<>> (update_body_relation.hpp:36-164) and
// UpdateRelation<P,Contact<>> (:166-288).  Gather form: each source particle
// collects its own neighbours; the inner pair criterion is symmetric, so the
// sets equal the reference's one-sided (i<j) atomic fill.
template <class R, int D>
static void update_relation(sphx_relation &rel)
{
    Body<R, D> &src = as<R, D>(rel.source);
    Body<R, D> &tar = as<R, D>(rel.target);
    CLL<R, D> &c = as<R, D>(rel.cll);
    NeighborKernel<R, D> nk(*src.kernel, src.h, tar.h);
    UInt n = src.n;
    rel.neighbor_size.assign(n, 0);
    rel.particle_offset.assign(n + 1, 0);
    auto visit = [&](UInt i, auto &&emit) {
        for_each_candidate(c, src.pos[i], [&](UInt j) {
            if (rel.inner && j == i)
                return;
            if (nk.criterion(src.pos[i], tar.pos[j]))
                emit(j);
        });
    };
    parallel_for(n, [&](size_t i) {
        UInt cnt = 0;
        visit((UInt)i, [&](UInt) { ++cnt; });
        rel.neighbor_size[i] = cnt;
    });
    std::vector<UInt> tmp(rel.neighbor_size);
    tmp.push_back(0);
    UInt total = exclusive_scan(tmp.data(), rel.particle_offset.data(), n + 1);
    rel.neighbor_index.assign(total, 0);
    parallel_for(n, [&](size_t i) {
        UInt w = rel.particle_offset[i];
        visit((UInt)i, [&](UInt j) { rel.neighbor_index[w++] = j; });
    });
    rel.built = true;
}

template <class R>
struct RiemannParams
{
    int kind;
    Impedance<R> imp;
    // riemann_solver_ck.hpp:45-57; common_functors.h:82-94 (TruncatedLinear slope 3)
    R DissipativePJump(R u_jump) const
    {
        if (kind == SPHX_RIEMANN_NONE)
            return R(0);
        R limiter = R(1);
        if (kind == SPHX_RIEMANN_ACOUSTIC)
            limiter = std::min(R(3.0) * (imp.inv_c0_ave * std::max(u_jump, R(0))), R(1));
        return imp.rho0c0_geo_ave * u_jump * limiter;
    }
    R DissipativeUJump(R p_jump) const
    {
        if (kind == SPHX_RIEMANN_NONE)
            return R(0);
        return p_jump * imp.inv_rho0c0_ave;
    }
    // AverageP / AverageV, riemann_solver_ck.hpp:17-30
    R AverageP(R p_i, R p_j) const { return imp.inv_rho0c0_sum * (p_i * imp.rho0c0 + p_j * imp.rho0c0); }
    // the same with LinearCorrectionCK's matrix arguments (CorrectionDataType = Matd)
    template <int D>
    Mat<R, D> AverageP(const Mat<R, D> &p_i, const Mat<R, D> &p_j) const
    {
        return imp.inv_rho0c0_sum * (p_i * imp.rho0c0 + p_j * imp.rho0c0);
    }
    template <int D>
    Vec<R, D> AverageV(const Vec<R, D> &vi, const Vec<R, D> &vj) const
    {
        return (vi * imp.rho0c0 + vj * imp.rho0c0) * imp.inv_rho0c0_sum;
    }
};

// corrected: LinearCorrectionCK (correction_(i) = B_i), else NoKernelCorrectionCK (1.0)
template <class R, int D>
static sphx_status acoustic_1st(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                                int riemann, double dt_d, bool corrected)
{
    Body<R, D> &b = as<R, D>(inner->source);
    const R dt = R(dt_d);
    const R rho0 = R(b.desc.rho0), c0 = R(b.desc.c0);
    const R p0 = rho0 * c0 * c0; // weakly_compressible_fluid.cpp:8-12
    RiemannParams<R> rs{riemann, Impedance<R>(rho0, c0)};
    // InitializeKernel::initialize, acoustic_step_1st_half.hpp:66-74
    parallel_for(b.n, [&](size_t i) {
        b.C[i] += 0.5 * dt * b.dC[i];
        b.rho[i] = b.C[i] * rho0;
        b.p[i] = p0 * (b.rho[i] / rho0 - 1.0);
        b.dpos[i] += b.vel[i] * dt * R(0.5);
    });
    NeighborKernel<R, D> nk(*b.kernel, b.h, b.h);
    // InteractKernel::interact (inner), acoustic_step_1st_half.hpp:89-111
    parallel_for(b.n, [&](size_t i) {
        Vec<R, D> force_sum = Vec<R, D>::Zero();
        R compression_dissipation(0);
        for (UInt n = inner->particle_offset[i]; n != inner->particle_offset[i + 1]; ++n)
        {
            UInt j = inner->neighbor_index[n];
            Vec<R, D> r = b.pos[i] - b.pos[j];
            R dW_ijV_j = nk.dW(r) * b.Vol[j];
            Vec<R, D> e_ij = normalized(r);
            if (corrected)
                force_sum -= rs.AverageP(b.correction[j] * b.p[i], b.correction[i] * b.p[j]) * R(2.0) *
                             dW_ijV_j * e_ij;
            else
                force_sum -= rs.AverageP(b.p[i], b.p[j]) * R(2.0) * dW_ijV_j * e_ij;
            compression_dissipation += rs.DissipativeUJump(b.p[i] - b.p[j]) * dW_ijV_j;
        }
        b.force[i] += force_sum * b.Vol[i];
        b.dC[i] = compression_dissipation * b.C[i];
    });
    // Contact<Wall> interact, acoustic_step_1st_half.hpp:157-180
    for (int k = 0; k < n_wall; ++k)
    {
        sphx_relation *w = walls[k];
        Body<R, D> &wb = as<R, D>(w->target);
        NeighborKernel<R, D> wk(*b.kernel, b.h, wb.h);
        parallel_for(b.n, [&](size_t i) {
            Vec<R, D> force_sum = Vec<R, D>::Zero();
            R compression_dissipation(0);
            for (UInt n = w->particle_offset[i]; n != w->particle_offset[i + 1]; ++n)
            {
                UInt j = w->neighbor_index[n];
                Vec<R, D> r = b.pos[i] - wb.pos[j];
                R dW_ijV_j = wk.dW(r) * wb.Vol[j];
                Vec<R, D> e_ij = normalized(r);
                R r_ij = norm(r);
                Vec<R, D> neg_e = e_ij * R(-1);
                R face_wall_external_acceleration = dot(b.force_prior[i] / b.mass[i] - wb.wall_acc[j], neg_e);
                R p_j_in_wall = b.p[i] + b.rho[i] * r_ij * std::max(R(0), face_wall_external_acceleration);
                if (corrected)
                    force_sum -= (b.p[i] + p_j_in_wall) * b.correction[i] * dW_ijV_j * e_ij;
                else
                    force_sum -= (b.p[i] + p_j_in_wall) * dW_ijV_j * e_ij;
                compression_dissipation += rs.DissipativeUJump(b.p[i] - p_j_in_wall) * dW_ijV_j;
            }
            b.force[i] += force_sum * b.Vol[i];
            b.dC[i] += compression_dissipation * b.C[i];
        });
    }
    // UpdateKernel::update, acoustic_step_1st_half.hpp:122-127
    parallel_for(b.n, [&](size_t i) {
        b.vel[i] += (b.force_prior[i] + b.force[i]) / b.mass[i] * dt;
    });
    return SPHX_OK;
}

template <class R, int D>
static sphx_status acoustic_2nd(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                                int riemann, double dt_d, bool corrected)
{
    Body<R, D> &b = as<R, D>(inner->source);
    const R dt = R(dt_d);
    const R rho0 = R(b.desc.rho0), c0 = R(b.desc.c0);
    RiemannParams<R> rs{riemann, Impedance<R>(rho0, c0)};
    // InitializeKernel::initialize, acoustic_step_2nd_half.hpp:33-38
    parallel_for(b.n, [&](size_t i) { b.dpos[i] += b.vel[i] * dt * R(0.5); });
    NeighborKernel<R, D> nk(*b.kernel, b.h, b.h);
    // inner interact, acoustic_step_2nd_half.hpp:53-73
    parallel_for(b.n, [&](size_t i) {
        R divergence_sum(0);
        Vec<R, D> p_dissipation = Vec<R, D>::Zero();
        for (UInt n = inner->particle_offset[i]; n != inner->particle_offset[i + 1]; ++n)
        {
            UInt j = inner->neighbor_index[n];
            Vec<R, D> r = b.pos[i] - b.pos[j];
            R dW_ijV_j = nk.dW(r) * b.Vol[j];
            Vec<R, D> e_ij = normalized(r);
            Vec<R, D> vel_ave = rs.AverageV(b.vel[i], b.vel[j]);
            const Vec<R, D> Be = corrected ? b.correction[i] * e_ij : e_ij;
            divergence_sum += R(2.0) * dot(b.vel[i] - vel_ave, Be) * dW_ijV_j;
            R u_jump = dot(b.vel[i] - b.vel[j], e_ij);
            p_dissipation += rs.DissipativePJump(u_jump) * dW_ijV_j * e_ij;
        }
        b.dC[i] += divergence_sum * b.C[i];
        b.force[i] = p_dissipation * b.Vol[i];
    });
    // Contact<Wall> interact, acoustic_step_2nd_half.hpp:117-137
    for (int k = 0; k < n_wall; ++k)
    {
        sphx_relation *w = walls[k];
        Body<R, D> &wb = as<R, D>(w->target);
        NeighborKernel<R, D> wk(*b.kernel, b.h, wb.h);
        parallel_for(b.n, [&](size_t i) {
            R divergence_sum(0);
            Vec<R, D> p_dissipation = Vec<R, D>::Zero();
            for (UInt n = w->particle_offset[i]; n != w->particle_offset[i + 1]; ++n)
            {
                UInt j = w->neighbor_index[n];
                Vec<R, D> r = b.pos[i] - wb.pos[j];
                R dW_ijV_j = wk.dW(r) * wb.Vol[j];
                Vec<R, D> e_ij = normalized(r);
                Vec<R, D> vel_diff = R(2.0) * (b.vel[i] - wb.wall_vel[j]);
                const Vec<R, D> Be = corrected ? b.correction[i] * e_ij : e_ij;
                divergence_sum += dot(vel_diff, Be) * dW_ijV_j;
                Vec<R, D> face_to_fluid_n = SGN(dot(e_ij, wb.normal[j])) * wb.normal[j];
                R u_jump = dot(vel_diff, face_to_fluid_n);
                p_dissipation += rs.DissipativePJump(u_jump) * dW_ijV_j * face_to_fluid_n;
            }
            b.dC[i] += divergence_sum * b.C[i];
            b.force[i] += p_dissipation * b.Vol[i];
        });
    }
    // UpdateKernel::update, acoustic_step_2nd_half.hpp:83-89
    parallel_for(b.n, [&](size_t i) {
        b.C[i] += 0.5 * dt * b.dC[i];
        b.rho[i] = b.C[i] * rho0;
    });
    return SPHX_OK;
}

// CompressionSummation<Inner<>> / <Contact<>>, density_regularization.hpp:40-83
template <class R, int D>
static sphx_status summation(sphx_relation *inner, sphx_relation *const *contact, int n_contact)
{
    Body<R, D> &b = as<R, D>(inner->source);
    NeighborKernel<R, D> nk(*b.kernel, b.h, b.h);
    parallel_for(b.n, [&](size_t i) {
        b.S[i] = nk.W0() * b.Vol_ref[i];
        for (UInt n = inner->particle_offset[i]; n != inner->particle_offset[i + 1]; ++n)
        {
            UInt j = inner->neighbor_index[n];
            b.S[i] += nk.W(b.pos[i] - b.pos[j]) * b.Vol_ref[j];
        }
    });
    for (int k = 0; k < n_contact; ++k)
    {
        sphx_relation *w = contact[k];
        Body<R, D> &wb = as<R, D>(w->target);
        NeighborKernel<R, D> wk(*b.kernel, b.h, wb.h);
        parallel_for(b.n, [&](size_t i) {
            for (UInt n = w->particle_offset[i]; n != w->particle_offset[i + 1]; ++n)
            {
                UInt j = w->neighbor_index[n];
                b.S[i] += wk.W(b.pos[i] - wb.pos[j]) * wb.Vol_ref[j];
            }
        });
    }
    return SPHX_OK;
}

// LinearCorrectionMatrix<Inner<WithUpdate>> + <Contact<>> (kernel_correction_ck.hpp:37-91):
// inner interact (B = 0, B -= r_ij grad W_ij^T V_j), contact interact, then the update
template <class R, int D>
static sphx_status correction_matrix(sphx_relation *inner, sphx_relation *const *contact, int n_contact,
                                     double alpha_d)
{
    Body<R, D> &b = as<R, D>(inner->source);
    const R alpha = R(alpha_d);
    NeighborKernel<R, D> nk(*b.kernel, b.h, b.h);
    // neighbor_method.hpp:88-94 (nablaW_ij = dW(r) r / |r|), kernel_correction_ck.hpp:39-49
    parallel_for(b.n, [&](size_t i) {
        Mat<R, D> Bi = Mat<R, D>::Zero();
        for (UInt n = inner->particle_offset[i]; n != inner->particle_offset[i + 1]; ++n)
        {
            UInt j = inner->neighbor_index[n];
            Vec<R, D> r = b.pos[i] - b.pos[j];
            Vec<R, D> nablaW_ijV_j = nk.dW(r) * normalized(r) * b.Vol[j];
            Bi -= outer(r, nablaW_ijV_j);
        }
        b.correction[i] = Bi;
    });
    // kernel_correction_ck.hpp:82-91
    for (int k = 0; k < n_contact; ++k)
    {
        sphx_relation *w = contact[k];
        Body<R, D> &wb = as<R, D>(w->target);
        NeighborKernel<R, D> wk(*b.kernel, b.h, wb.h);
        parallel_for(b.n, [&](size_t i) {
            for (UInt n = w->particle_offset[i]; n != w->particle_offset[i + 1]; ++n)
            {
                UInt j = w->neighbor_index[n];
                Vec<R, D> r = b.pos[i] - wb.pos[j];
                Vec<R, D> nablaW_ijV_j = wk.dW(r) * normalized(r) * wb.Vol[j];
                b.correction[i] -= outer(r, nablaW_ijV_j);
            }
        });
    }
    // UpdateKernel::update, kernel_correction_ck.hpp:58-65
    parallel_for(b.n, [&](size_t i) {
        const Mat<R, D> &Bi = b.correction[i];
        R determinant_i = determinant(Bi);
        R det_sqr = std::max(alpha - determinant_i, R(0));
        Mat<R, D> inv = inverseTikhonov(Bi, SqrtEps<R>());
        R weight = determinant_i / (determinant_i + det_sqr);
        b.correction[i] = weight * inv + (R(1.0) - weight) * Mat<R, D>::Identity();
    });
    return SPHX_OK;
}

// FreeSurfaceIndicationCK<Inner<WithUpdate>> + <Contact<>> (surface_indication_ck.hpp:11-178)
template <class R, int D>
static sphx_status surface_indication(sphx_relation *inner, sphx_relation *const *contact, int n_contact)
{
    Body<R, D> &b = as<R, D>(inner->source);
    const R threshold = R(0.75) * R(D); // surface_indication_ck.hpp:19
    const R smoothing_length = b.h;     // ReferenceSmoothingLength, :20
    NeighborKernel<R, D> nk(*b.kernel, b.h, b.h);
    auto first = [&](size_t i) { return inner->particle_offset[i]; };
    auto last = [&](size_t i) { return inner->particle_offset[i + 1]; };
    // inner InteractKernel::interact, :55-71, with isNearPreviousFreeSurface, :75-89
    parallel_for(b.n, [&](size_t i) {
        R pos_div = 0.0;
        for (UInt n = first(i); n != last(i); ++n)
        {
            UInt j = inner->neighbor_index[n];
            R r_ij = norm(b.pos[i] - b.pos[j]);
            pos_div -= nk.dW(b.pos[i] - b.pos[j]) * b.Vol[j] * r_ij;
        }
        auto near_previous_surface = [&]() {
            for (UInt n = first(i); n != last(i); ++n)
                if (b.prev_indicator[inner->neighbor_index[n]] == 1)
                    return true;
            return false;
        };
        if (pos_div < threshold && b.prev_indicator[i] != 1 && !near_previous_surface())
            pos_div = R(2.0) * threshold;
        b.pos_div[i] = pos_div;
    });
    // contact InteractKernel::interact, :147-157
    for (int k = 0; k < n_contact; ++k)
    {
        sphx_relation *w = contact[k];
        Body<R, D> &wb = as<R, D>(w->target);
        NeighborKernel<R, D> wk(*b.kernel, b.h, wb.h);
        parallel_for(b.n, [&](size_t i) {
            R pos_div = 0.0;
            for (UInt n = w->particle_offset[i]; n != w->particle_offset[i + 1]; ++n)
            {
                UInt j = w->neighbor_index[n];
                R r_ij = norm(b.pos[i] - wb.pos[j]);
                pos_div -= wk.dW(b.pos[i] - wb.pos[j]) * wb.Vol[j] * r_ij;
            }
            b.pos_div[i] += pos_div;
        });
    }
    // UpdateKernel::update, :100-109, with isVeryNearFreeSurface, :112-127
    parallel_for(b.n, [&](size_t i) {
        auto very_near_surface = [&]() {
            for (UInt n = first(i); n != last(i); ++n)
            {
                UInt j = inner->neighbor_index[n];
                R r_ij = norm(b.pos[i] - b.pos[j]);
                if (b.pos_div[j] < threshold && r_ij < smoothing_length)
                    return true;
            }
            return false;
        };
        b.indicator[i] = 1;
        if (b.pos_div[i] > threshold && !very_near_surface())
            b.indicator[i] = 0;
        b.prev_indicator[i] = b.indicator[i];
    });
    return SPHX_OK;
}

// ViscousForceCK<Inner<WithUpdate, Viscosity, K>> + <Contact<Wall, Viscosity, K>>
// (viscous_force.hpp:14-104) and the ForcePriorCK update (force_prior_ck.h:55-59)
template <class R, int D>
static sphx_status viscous_force(sphx_relation *inner, sphx_relation *const *walls, int n_wall, double mu_d,
                                 bool corrected)
{
    Body<R, D> &b = as<R, D>(inner->source);
    const R mu = R(mu_d);
    // Viscosity::getInterParticleViscosity = PairGeomAverageFixed(mu, mu), particle_functors.h:169-170
    const R mu_pair = R(2.0) * mu * mu * (R(1) / (mu + mu));
    const R h_sq = std::pow(b.h, R(2)); // pow(ReferenceSmoothingLength(), 2), viscous_force.hpp:24
    NeighborKernel<R, D> nk(*b.kernel, b.h, b.h);
    parallel_for(b.n, [&](size_t i) {
        Vec<R, D> force = Vec<R, D>::Zero();
        for (UInt n = inner->particle_offset[i]; n != inner->particle_offset[i + 1]; ++n)
        {
            UInt j = inner->neighbor_index[n];
            Vec<R, D> r = b.pos[i] - b.pos[j];
            Vec<R, D> e_ij = normalized(r);
            R dW_ijV_j = nk.dW(r) * b.Vol[j];
            Vec<R, D> vel_derivative = (b.vel[i] - b.vel[j]) / (squaredNorm(r) + R(0.01) * h_sq);
            Vec<R, D> ce = corrected ? (b.correction[i] + b.correction[j]) * e_ij : (R(1.0) + R(1.0)) * e_ij;
            force += dot(r, ce) * mu_pair * vel_derivative * dW_ijV_j;
        }
        b.viscous_force[i] = force * b.Vol[i];
    });
    for (int k = 0; k < n_wall; ++k)
    {
        sphx_relation *w = walls[k];
        Body<R, D> &wb = as<R, D>(w->target);
        NeighborKernel<R, D> wk(*b.kernel, b.h, wb.h);
        parallel_for(b.n, [&](size_t i) {
            Vec<R, D> force = Vec<R, D>::Zero();
            for (UInt n = w->particle_offset[i]; n != w->particle_offset[i + 1]; ++n)
            {
                UInt j = w->neighbor_index[n];
                Vec<R, D> r = b.pos[i] - wb.pos[j];
                Vec<R, D> e_ij = normalized(r);
                R dW_ijV_j = wk.dW(r) * wb.Vol[j];
                Vec<R, D> vel_derivative = R(2.0) * (b.vel[i] - wb.wall_vel[j]) / (squaredNorm(r) + R(0.01) * h_sq);
                Vec<R, D> ce = corrected ? b.correction[i] * e_ij : R(1.0) * e_ij;
                force += R(2.0) * dot(r, ce) * mu * vel_derivative * dW_ijV_j;
            }
            b.viscous_force[i] += force * b.Vol[i];
        });
    }
    parallel_for(b.n, [&](size_t i) {
        b.force_prior[i] += b.viscous_force[i] - b.prev_viscous_force[i];
        b.prev_viscous_force[i] = b.viscous_force[i];
    });
    return SPHX_OK;
}

// KernelGradientIntegral<Inner<K>> + <Contact<Boundary, K>> (kernel_gradient_integral.hpp:37-49,72-84)
template <class R, int D>
static sphx_status kernel_gradient_integral(sphx_relation *inner, sphx_relation *const *contact, int n_contact,
                                            bool corrected)
{
    Body<R, D> &b = as<R, D>(inner->source);
    NeighborKernel<R, D> nk(*b.kernel, b.h, b.h);
    parallel_for(b.n, [&](size_t i) {
        Vec<R, D> inconsistency = Vec<R, D>::Zero();
        for (UInt n = inner->particle_offset[i]; n != inner->particle_offset[i + 1]; ++n)
        {
            UInt j = inner->neighbor_index[n];
            Vec<R, D> r = b.pos[i] - b.pos[j];
            const R dW_ijV_j = nk.dW(r) * b.Vol[j];
            const Vec<R, D> e_ij = normalized(r);
            if (corrected)
                inconsistency -= (b.correction[i] + b.correction[j]) * dW_ijV_j * e_ij;
            else
                inconsistency -= (R(1.0) + R(1.0)) * dW_ijV_j * e_ij;
        }
        b.kernel_gradient_integral[i] = inconsistency;
    });
    for (int k = 0; k < n_contact; ++k)
    {
        sphx_relation *w = contact[k];
        Body<R, D> &wb = as<R, D>(w->target);
        NeighborKernel<R, D> wk(*b.kernel, b.h, wb.h);
        parallel_for(b.n, [&](size_t i) {
            Vec<R, D> inconsistency = Vec<R, D>::Zero();
            for (UInt n = w->particle_offset[i]; n != w->particle_offset[i + 1]; ++n)
            {
                UInt j = w->neighbor_index[n];
                Vec<R, D> r = b.pos[i] - wb.pos[j];
                const R dW_ijV_j = wk.dW(r) * wb.Vol[j];
                const Vec<R, D> e_ij = normalized(r);
                if (corrected)
                    inconsistency -= R(2.0) * b.correction[i] * dW_ijV_j * e_ij;
                else
                    inconsistency -= R(2.0) * R(1.0) * dW_ijV_j * e_ij;
            }
            b.kernel_gradient_integral[i] += inconsistency;
        });
    }
    return SPHX_OK;
}

// fixed-block summation so that the result does not depend on the thread count
template <class F>
static double block_sum(size_t n, const F &f)
{
    const size_t B = 4096;
    size_t nb = (n + B - 1) / B;
    std::vector<double> part(nb, 0.0);
    parallel_for(nb, [&](size_t b) {
        double s = 0;
        for (size_t i = b * B; i < std::min(n, (b + 1) * B); ++i)
            s += f(i);
        part[b] = s;
    });
    double s = 0;
    for (double v : part)
        s += v;
    return s;
}
template <class R, class F>
static R block_max(size_t n, const F &f)
{
    const size_t B = 4096;
    size_t nb = (n + B - 1) / B;
    std::vector<R> part(nb, std::numeric_limits<R>::lowest());
    parallel_for(nb, [&](size_t b) {
        R m = std::numeric_limits<R>::lowest();
        for (size_t i = b * B; i < std::min(n, (b + 1) * B); ++i)
            m = std::max(m, f(i));
        part[b] = m;
    });
    R m = std::numeric_limits<R>::lowest(); // ReduceReference<ReduceMax> = MinReal
    for (R v : part)
        m = std::max(m, v);
    return m;
}

//=============================================================================
// C ABI
//=============================================================================
template <class R, int D>
static sphx_status rw_vec(std::vector<Vec<R, D>> &f, void *host, size_t n, bool write)
{
    double *h = (double *)host;
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < D; ++k)
        {
            if (write)
                f[i][k] = R(h[i * D + k]);
            else
                h[i * D + k] = (double)f[i][k];
        }
    return SPHX_OK;
}
template <class R>
static sphx_status rw_scalar(std::vector<R> &f, void *host, size_t n, bool write)
{
    double *h = (double *)host;
    for (size_t i = 0; i < n; ++i)
    {
        if (write)
            f[i] = R(h[i]);
        else
            h[i] = (double)f[i];
    }
    return SPHX_OK;
}
template <class R, int D>
static sphx_status rw_mat(std::vector<Mat<R, D>> &f, void *host, size_t n, bool write)
{
    double *h = (double *)host;
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b)
            {
                if (write)
                    f[i].m[a][b] = R(h[(i * D + a) * D + b]);
                else
                    h[(i * D + a) * D + b] = (double)f[i].m[a][b];
            }
    return SPHX_OK;
}
static sphx_status rw_int(std::vector<int32_t> &f, void *host, size_t n, bool write)
{
    if (write)
        std::memcpy(f.data(), host, n * sizeof(int32_t));
    else
        std::memcpy(host, f.data(), n * sizeof(int32_t));
    return SPHX_OK;
}
static sphx_status rw_uint(std::vector<UInt> &f, void *host, size_t n, bool write)
{
    if (write)
        std::memcpy(f.data(), host, n * sizeof(UInt));
    else
        std::memcpy(host, f.data(), n * sizeof(UInt));
    return SPHX_OK;
}

extern "C" {

const char *sphx_backend_name(void) { return "oracle-cpu"; }
const char *sphx_last_error(void) { return g_last_error.c_str(); }

sphx_status sphx_context_create(int device, sphx_context **out)
{
    if (!out)
        return fail(SPHX_E_ARG, "null out");
    *out = new sphx_context{device};
    return SPHX_OK;
}
sphx_status sphx_context_destroy(sphx_context *ctx)
{
    delete ctx;
    return SPHX_OK;
}
sphx_status sphx_context_synchronize(sphx_context *) { return SPHX_OK; }

sphx_status sphx_body_create(sphx_context *ctx, const sphx_body_desc *d, sphx_body **out)
{
    if (!ctx || !d || !out)
        return fail(SPHX_E_ARG, "null argument");
    if (d->dim != 2 && d->dim != 3)
        return fail(SPHX_E_ARG, "dim must be 2 or 3");
    if (d->precision != SPHX_FP64 && d->precision != ORACLE_FP32)
        return fail(SPHX_E_ARG, "oracle supports FP64 and FP32 only");
    if (!(d->spacing_ref > 0) || !(d->h_spacing_ratio > 0))
        return fail(SPHX_E_ARG, "spacing_ref and h_spacing_ratio must be positive");
    auto *b = new sphx_body{ctx, nullptr};
    if (d->precision == SPHX_FP64)
        b->impl.reset(d->dim == 2 ? (BodyBase *)new Body<double, 2>(*d) : new Body<double, 3>(*d));
    else
        b->impl.reset(d->dim == 2 ? (BodyBase *)new Body<float, 2>(*d) : new Body<float, 3>(*d));
    *out = b;
    return SPHX_OK;
}
sphx_status sphx_body_destroy(sphx_body *b)
{
    delete b;
    return SPHX_OK;
}

sphx_status sphx_body_smoothing_length(const sphx_body *b, double *h)
{
    return with_body(const_cast<sphx_body *>(b), [&](auto &B) {
        *h = (double)B.h;
        return SPHX_OK;
    });
}

static sphx_status body_rw(sphx_body *b, int field, void *host, size_t n, bool write)
{
    if (!host)
        return fail(SPHX_E_ARG, "null host pointer");
    return with_body(b, [&](auto &B) -> sphx_status {
        if (n != B.n)
            return fail(SPHX_E_ARG, "particle count mismatch");
        switch (field)
        {
        case SPHX_FIELD_POSITION: return rw_vec(B.pos, host, n, write);
        case SPHX_FIELD_VOLUME:
        {
            sphx_status s = rw_scalar(B.Vol, host, n, write);
            if (write && !B.vol_written)
            {
                // first write: VolumetricMeasureRef from VolumetricMeasure
                // (density_regularization.hpp:18-20) and Mass = rho0 * Vol
                // (base_material.cpp:37-40)
                B.vol_written = true;
                using R = typename std::decay<decltype(B)>::type::Real;
                for (size_t i = 0; i < n; ++i)
                {
                    B.Vol_ref[i] = B.Vol[i];
                    B.mass[i] = R(B.desc.rho0) * B.Vol[i];
                }
            }
            return s;
        }
        case SPHX_FIELD_MASS: return rw_scalar(B.mass, host, n, write);
        case SPHX_FIELD_DENSITY: return rw_scalar(B.rho, host, n, write);
        case SPHX_FIELD_PRESSURE: return rw_scalar(B.p, host, n, write);
        case SPHX_FIELD_VELOCITY: return rw_vec(B.vel, host, n, write);
        case SPHX_FIELD_FORCE: return rw_vec(B.force, host, n, write);
        case SPHX_FIELD_FORCE_PRIOR: return rw_vec(B.force_prior, host, n, write);
        case SPHX_FIELD_DISPLACEMENT: return rw_vec(B.dpos, host, n, write);
        case SPHX_FIELD_COMPRESSION: return rw_scalar(B.C, host, n, write);
        case SPHX_FIELD_COMPRESSION_RATE: return rw_scalar(B.dC, host, n, write);
        case SPHX_FIELD_VOLUME_REF: return rw_scalar(B.Vol_ref, host, n, write);
        case SPHX_FIELD_COMPRESSION_SUM: return rw_scalar(B.S, host, n, write);
        case SPHX_FIELD_NORMAL: return rw_vec(B.normal, host, n, write);
        case SPHX_FIELD_WALL_VELOCITY: return rw_vec(B.wall_vel, host, n, write);
        case SPHX_FIELD_WALL_ACCELERATION: return rw_vec(B.wall_acc, host, n, write);
        case SPHX_FIELD_ORIGINAL_ID: return rw_uint(B.original_id, host, n, write);
        case SPHX_FIELD_SORTED_ID: return rw_uint(B.sorted_id, host, n, write);
        case SPHX_FIELD_SORT_SEQUENCE: return rw_uint(B.sequence, host, n, write);
        case SPHX_FIELD_INTERPOLATED: return rw_scalar(B.interp, host, n, write);
        case SPHX_FIELD_CORRECTION_MATRIX: return rw_mat(B.correction, host, n, write);
        case SPHX_FIELD_INDICATOR: return rw_int(B.indicator, host, n, write);
        case SPHX_FIELD_PREVIOUS_INDICATOR: return rw_int(B.prev_indicator, host, n, write);
        case SPHX_FIELD_POSITION_DIVERGENCE: return rw_scalar(B.pos_div, host, n, write);
        case SPHX_FIELD_VISCOUS_FORCE: return rw_vec(B.viscous_force, host, n, write);
        case SPHX_FIELD_PREVIOUS_VISCOUS_FORCE: return rw_vec(B.prev_viscous_force, host, n, write);
        case SPHX_FIELD_KERNEL_GRADIENT_INTEGRAL: return rw_vec(B.kernel_gradient_integral, host, n, write);
        default: return fail(SPHX_E_ARG, "unknown field");
        }
    });
}
sphx_status sphx_body_write(sphx_body *b, int field, const void *host, size_t n)
{
    return body_rw(b, field, const_cast<void *>(host), n, true);
}
sphx_status sphx_body_read(sphx_body *b, int field, void *host, size_t n)
{
    return body_rw(b, field, host, n, false);
}

sphx_status sphx_cell_linked_list_create(sphx_body *body, const double *lo, const double *up,
                                         double gs, uint32_t buffer_width,
                                         sphx_cell_linked_list **out)
{
    if (!body || !lo || !up || !out || !(gs > 0))
        return fail(SPHX_E_ARG, "bad cell linked list arguments");
    auto *c = new sphx_cell_linked_list{body, nullptr};
    with_body(body, [&](auto &B) {
        BODY_TYPES(B);
        auto *impl = new CLL<R, D>();
        Vec<R, D> bl, bu;
        for (int k = 0; k < D; ++k)
            bl[k] = R(lo[k]), bu[k] = R(up[k]);
        impl->mesh = Mesh<R, D>(bl, bu, R(gs), buffer_width);
        c->impl.reset(impl);
        return SPHX_OK;
    });
    *out = c;
    return SPHX_OK;
}
sphx_status sphx_cell_linked_list_destroy(sphx_cell_linked_list *c)
{
    delete c;
    return SPHX_OK;
}
sphx_status sphx_cell_linked_list_mesh(const sphx_cell_linked_list *c, int *all_cells,
                                       double *lower, double *gs)
{
    auto *cc = const_cast<sphx_cell_linked_list *>(c);
    return with_body(cc->body, [&](auto &B) {
        BODY_TYPES(B);
        CLL<R, D> &L = as<R, D>(cc);
        for (int k = 0; k < D; ++k)
        {
            if (all_cells)
                all_cells[k] = L.mesh.all_cells[k];
            if (lower)
                lower[k] = (double)L.mesh.lower[k];
        }
        if (gs)
            *gs = (double)L.mesh.gs;
        return SPHX_OK;
    });
}
sphx_status sphx_update_cell_linked_list(sphx_cell_linked_list *c)
{
    if (!c)
        return fail(SPHX_E_ARG, "null cll");
    return with_body(c->body, [&](auto &B) {
        BODY_TYPES(B);
        update_cll(B, as<R, D>(c));
        return SPHX_OK;
    });
}
// the CPU restatement holds no device workspace: the same update
sphx_status sphx_update_cell_linked_list_noalloc(sphx_cell_linked_list *c) { return sphx_update_cell_linked_list(c); }
sphx_status sphx_cell_linked_list_read(sphx_cell_linked_list *c, uint32_t *cell_offset, size_t n_off,
                                       uint32_t *particle_index, size_t n_idx)
{
    if (!c)
        return fail(SPHX_E_ARG, "null cll");
    return with_body(c->body, [&](auto &B) -> sphx_status {
        BODY_TYPES(B);
        CLL<R, D> &L = as<R, D>(c);
        if (!L.built)
            return fail(SPHX_E_STATE, "cell linked list never updated");
        if (cell_offset)
        {
            if (n_off != L.cell_offset.size())
                return fail(SPHX_E_CAPACITY, "cell_offset size must be n_cells+1");
            std::memcpy(cell_offset, L.cell_offset.data(), n_off * sizeof(UInt));
        }
        if (particle_index)
        {
            if (n_idx != L.particle_index.size())
                return fail(SPHX_E_CAPACITY, "particle_index size must be n_particles");
            std::memcpy(particle_index, L.particle_index.data(), n_idx * sizeof(UInt));
        }
        return SPHX_OK;
    });
}

sphx_status sphx_relation_create_inner(sphx_body *body, sphx_cell_linked_list *cll, sphx_relation **out)
{
    if (!body || !cll || !out || cll->body != body)
        return fail(SPHX_E_ARG, "inner relation needs the body's own cell linked list");
    *out = new sphx_relation{true, body, body, cll, {}, {}, {}, false};
    return SPHX_OK;
}
sphx_status sphx_relation_create_contact(sphx_body *src, sphx_body *tar, sphx_cell_linked_list *tcll,
                                         sphx_relation **out)
{
    if (!src || !tar || !tcll || !out || tcll->body != tar || !same_kind(src, tar))
        return fail(SPHX_E_ARG, "contact relation needs the target's cell linked list");
    *out = new sphx_relation{false, src, tar, tcll, {}, {}, {}, false};
    return SPHX_OK;
}
sphx_status sphx_relation_destroy(sphx_relation *r)
{
    delete r;
    return SPHX_OK;
}
sphx_status sphx_update_relation(sphx_relation *r)
{
    if (!r)
        return fail(SPHX_E_ARG, "null relation");
    return with_body(r->source, [&](auto &B) -> sphx_status {
        BODY_TYPES(B);
        if (!as<R, D>(r->cll).built)
            return fail(SPHX_E_STATE, "target cell linked list never updated");
        update_relation<R, D>(*r);
        return SPHX_OK;
    });
}
sphx_status sphx_relation_read(sphx_relation *r, uint32_t *size, uint32_t *offset, uint32_t *index,
                               size_t capacity, size_t *total)
{
    if (!r)
        return fail(SPHX_E_ARG, "null relation");
    if (!r->built)
        return fail(SPHX_E_STATE, "relation never updated");
    if (total)
        *total = r->neighbor_index.size();
    if (size)
        std::memcpy(size, r->neighbor_size.data(), r->neighbor_size.size() * sizeof(UInt));
    if (offset)
        std::memcpy(offset, r->particle_offset.data(), r->particle_offset.size() * sizeof(UInt));
    if (index)
    {
        if (capacity < r->neighbor_index.size())
            return fail(SPHX_E_CAPACITY, "neighbor_index capacity too small");
        std::memcpy(index, r->neighbor_index.data(), r->neighbor_index.size() * sizeof(UInt));
    }
    return SPHX_OK;
}

// ParticleSortCK<P>::exec, particle_sort_ck.hpp:61-104; sort-sequence mesh is
// the CLL mesh with linear offset 0 (cell_linked_list.cpp:131-136)
sphx_status sphx_particle_sort(sphx_body *body, sphx_cell_linked_list *cll)
{
    if (!body || !cll || cll->body != body)
        return fail(SPHX_E_ARG, "sort needs the body's own cell linked list");
    return with_body(body, [&](auto &B) {
        BODY_TYPES(B);
        CLL<R, D> &L = as<R, D>(cll);
        std::vector<UInt> perm(B.n);
        for (UInt i = 0; i < B.n; ++i)
        {
            int c[D];
            L.mesh.CellIndexFromPosition(B.pos[i], c);
            B.sequence[i] = Mesh<R, D>::MortonOrder(c);
            perm[i] = i;
        }
        std::stable_sort(perm.begin(), perm.end(),
                         [&](UInt a, UInt b) { return B.sequence[a] < B.sequence[b]; });
        auto gather = [&](auto &v) {
            auto tmp = v;
            for (UInt i = 0; i < B.n; ++i)
                v[i] = tmp[perm[i]];
        };
        gather(B.sequence);
        // evolving variables only (base_particles.cpp:33-37, acoustic_step_1st_half.hpp:30-34,
        // force_prior_ck.cpp:13, density_regularization.hpp:24)
        gather(B.pos), gather(B.Vol), gather(B.original_id), gather(B.vel), gather(B.mass);
        gather(B.force_prior), gather(B.C), gather(B.dC), gather(B.Vol_ref);
        // evolving once the free-surface indication exists (surface_indication_ck.hpp:43)
        gather(B.prev_indicator);
        // evolving with a viscous force (force_prior_ck.cpp:14)
        gather(B.prev_viscous_force);
        for (UInt i = 0; i < B.n; ++i)
            B.sorted_id[B.original_id[i]] = i;
        return SPHX_OK;
    });
}

// GravityForceCK: force_prior += m g - previous (previous = 0), force_prior_ck.hpp:37-43
sphx_status sphx_gravity_force(sphx_body *body, const double *g)
{
    if (!g)
        return fail(SPHX_E_ARG, "null gravity");
    return with_body(body, [&](auto &B) {
        BODY_TYPES(B);
        Vec<R, D> gv;
        for (int k = 0; k < D; ++k)
            gv[k] = R(g[k]);
        for (UInt i = 0; i < B.n; ++i)
            B.force_prior[i] += B.mass[i] * gv;
        return SPHX_OK;
    });
}
// density_regularization.hpp:108-118 with Regularization<FreeSurface|Internal> (.h:146-184)
sphx_status sphx_density_regularization(sphx_body *body, int flow)
{
    return with_body(body, [&](auto &B) {
        using R = typename std::decay<decltype(B)>::type::Real;
        R rho0 = R(B.desc.rho0);
        parallel_for(B.n, [&](size_t i) {
            B.C[i] = flow == SPHX_FLOW_FREE_SURFACE ? std::max(B.S[i], R(1)) : B.S[i];
            B.rho[i] = B.C[i] * rho0;
        });
        return SPHX_OK;
    });
}
// fluid_time_step_ck.h:148-152
sphx_status sphx_advection_step_setup(sphx_body *body)
{
    return with_body(body, [&](auto &B) {
        BODY_TYPES(B);
        parallel_for(B.n, [&](size_t i) {
            B.Vol[i] = B.mass[i] / B.rho[i];
            B.dpos[i] = Vec<R, D>::Zero();
        });
        return SPHX_OK;
    });
}
// fluid_time_step_ck.h:176-179
sphx_status sphx_update_particle_position(sphx_body *body)
{
    return with_body(body, [&](auto &B) {
        parallel_for(B.n, [&](size_t i) { B.pos[i] += B.dpos[i]; });
        return SPHX_OK;
    });
}

static bool relations_ok(sphx_relation *inner, sphx_relation *const *walls, int n_wall)
{
    if (!inner || !inner->inner || !inner->built || n_wall < 0 || (n_wall > 0 && !walls))
        return false;
    for (int k = 0; k < n_wall; ++k)
        if (!walls[k] || walls[k]->inner || walls[k]->source != inner->source || !walls[k]->built)
            return false;
    return true;
}

sphx_status sphx_compression_summation(sphx_relation *inner, sphx_relation *const *contact, int n)
{
    if (!relations_ok(inner, contact, n))
        return fail(SPHX_E_ARG, "summation needs an updated inner relation and contact relations of the same source");
    return with_body(inner->source, [&](auto &B) {
        BODY_TYPES(B);
        return summation<R, D>(inner, contact, n);
    });
}
static sphx_status first_half(sphx_relation *inner, sphx_relation *const *walls, int n, int riemann, double dt,
                              bool corrected)
{
    if (!relations_ok(inner, walls, n))
        return fail(SPHX_E_ARG, "acoustic step needs updated relations of one source");
    return with_body(inner->source, [&](auto &B) {
        BODY_TYPES(B);
        return acoustic_1st<R, D>(inner, walls, n, riemann, dt, corrected);
    });
}
static sphx_status second_half(sphx_relation *inner, sphx_relation *const *walls, int n, int riemann, double dt,
                               bool corrected)
{
    if (!relations_ok(inner, walls, n))
        return fail(SPHX_E_ARG, "acoustic step needs updated relations of one source");
    return with_body(inner->source, [&](auto &B) {
        BODY_TYPES(B);
        return acoustic_2nd<R, D>(inner, walls, n, riemann, dt, corrected);
    });
}
sphx_status sphx_acoustic_step_1st_half(sphx_relation *inner, sphx_relation *const *walls, int n,
                                        int riemann, double dt)
{
    return first_half(inner, walls, n, riemann, dt, false);
}
sphx_status sphx_acoustic_step_2nd_half(sphx_relation *inner, sphx_relation *const *walls, int n,
                                        int riemann, double dt)
{
    return second_half(inner, walls, n, riemann, dt, false);
}
// the device library's cell-traversal form has the neighbours of the lists: on
// the CPU it is the 2nd half itself
sphx_status sphx_acoustic_step_2nd_half_traversal(sphx_relation *inner, sphx_relation *const *walls, int n,
                                                  int riemann, double dt)
{
    return second_half(inner, walls, n, riemann, dt, false);
}
sphx_status sphx_acoustic_step_1st_half_corrected(sphx_relation *inner, sphx_relation *const *walls, int n,
                                                  int riemann, double dt)
{
    return first_half(inner, walls, n, riemann, dt, true);
}
sphx_status sphx_acoustic_step_2nd_half_corrected(sphx_relation *inner, sphx_relation *const *walls, int n,
                                                  int riemann, double dt)
{
    return second_half(inner, walls, n, riemann, dt, true);
}
sphx_status sphx_linear_correction_matrix(sphx_relation *inner, sphx_relation *const *contact, int n,
                                          double alpha)
{
    if (!relations_ok(inner, contact, n))
        return fail(SPHX_E_ARG, "correction matrix needs updated relations of one source");
    return with_body(inner->source, [&](auto &B) {
        BODY_TYPES(B);
        return correction_matrix<R, D>(inner, contact, n, alpha);
    });
}
sphx_status sphx_viscous_force(sphx_relation *inner, sphx_relation *const *walls, int n, double mu, int correction)
{
    if (!relations_ok(inner, walls, n) || !(mu >= 0.0))
        return fail(SPHX_E_ARG, "viscous force needs updated relations of one source and mu >= 0");
    return with_body(inner->source, [&](auto &B) {
        BODY_TYPES(B);
        return viscous_force<R, D>(inner, walls, n, mu, correction == SPHX_CORRECTION_LINEAR);
    });
}
sphx_status sphx_kernel_gradient_integral(sphx_relation *inner, sphx_relation *const *contact, int n, int correction)
{
    if (!relations_ok(inner, contact, n))
        return fail(SPHX_E_ARG, "kernel gradient integral needs updated relations of one source");
    return with_body(inner->source, [&](auto &B) {
        BODY_TYPES(B);
        return kernel_gradient_integral<R, D>(inner, contact, n, correction == SPHX_CORRECTION_LINEAR);
    });
}
// TransportVelocityCorrectionCK::UpdateKernel::update, transport_velocity_correction_ck.hpp:39-51;
// TruncatedLinear(100), common_functors.h:82-94; BulkParticles: Indicator == 0
sphx_status sphx_transport_velocity_correction(sphx_body *body, double coefficient, int limiter, int bulk_only)
{
    return with_body(body, [&](auto &B) {
        BODY_TYPES(B);
        const R h_ref = B.h;
        const R squared_h_ref = h_ref * h_ref;
        const R correction_scaling = R(coefficient) * squared_h_ref;
        parallel_for(B.n, [&](size_t i) {
            if (bulk_only && B.indicator[i] != 0)
                return;
            const R inv_h_ratio = R(1.0) / R(1.0);
            const Vec<R, D> &residual = B.kernel_gradient_integral[i];
            const R squared_norm = squaredNorm(residual);
            const R lim = limiter == SPHX_LIMITER_TRUNCATED_LINEAR
                              ? std::min(R(100.0) * (squared_h_ref * squared_norm), R(1))
                              : R(1.0);
            B.dpos[i] += correction_scaling * lim * B.kernel_gradient_integral[i] * inv_h_ratio * inv_h_ratio;
        });
        return SPHX_OK;
    });
}
sphx_status sphx_free_surface_indication(sphx_relation *inner, sphx_relation *const *contact, int n)
{
    if (!relations_ok(inner, contact, n))
        return fail(SPHX_E_ARG, "surface indication needs updated relations of one source");
    return with_body(inner->source, [&](auto &B) {
        BODY_TYPES(B);
        return surface_indication<R, D>(inner, contact, n);
    });
}

// Not in the reference (its max reductions drop NaN): a time step that is not
// finite and positive comes from a particle with a non-finite velocity, force
// or compression, and is reported instead of returned
static sphx_status finite_step(double dt, const char *which)
{
    if (std::isfinite(dt) && dt > 0.0)
        return SPHX_OK;
    return fail(SPHX_E_NONFINITE, std::string("non-finite ") + which +
                                      " time step: a particle's velocity, force or compression is not finite");
}

// AcousticTimeStepCK, fluid_time_step_ck.hpp:30-57 (h_min = reference smoothing length)
sphx_status sphx_acoustic_time_step(sphx_body *body, double cfl, double *dt)
{
    if (!dt)
        return fail(SPHX_E_ARG, "null dt");
    return with_body(body, [&](auto &B) {
        using R = typename std::decay<decltype(B)>::type::Real;
        R h_min = B.h, c0 = R(B.desc.c0);
        R m = block_max<R>(B.n, [&](size_t i) {
            R force_norm = norm(B.force[i] + B.force_prior[i]);
            R acc_scale = std::sqrt(4.0 * h_min * force_norm / B.mass[i]);
            R signal = c0 + norm(B.vel[i]);
            // a particle with a non-finite state takes +inf (the step becomes zero)
            if (!std::isfinite(signal) || !std::isfinite(acc_scale) || !std::isfinite(B.C[i]))
                return std::numeric_limits<R>::infinity();
            return std::max(signal, acc_scale);
        });
        *dt = (double)(R(cfl) * h_min / (m + TinyReal<R>()));
        return finite_step(*dt, "acoustic");
    });
}
// AdvectionTimeStepCK, fluid_time_step_ck.cpp:23-26, fluid_time_step_ck.h:107-110
sphx_status sphx_advection_time_step(sphx_body *body, double u_ref, double cfl, double *dt)
{
    if (!dt)
        return fail(SPHX_E_ARG, "null dt");
    return with_body(body, [&](auto &B) {
        using R = typename std::decay<decltype(B)>::type::Real;
        R m = block_max<R>(B.n, [&](size_t i) {
            R v2 = squaredNorm(B.vel[i]);
            return std::isfinite(v2) ? v2 : std::numeric_limits<R>::infinity();
        });
        *dt = (double)(R(cfl) * B.h / (std::max(R(std::sqrt(m)), R(u_ref)) + TinyReal<R>()));
        return finite_step(*dt, "advection");
    });
}
// TotalMechanicalEnergyCK, general_reduce_ck.h:86-90; Gravity::getPotential external_force.h:53-56
sphx_status sphx_total_mechanical_energy(sphx_body *body, const double *g, double *e)
{
    if (!g || !e)
        return fail(SPHX_E_ARG, "null argument");
    return with_body(body, [&](auto &B) {
        BODY_TYPES(B);
        Vec<R, D> gv;
        for (int k = 0; k < D; ++k)
            gv[k] = R(g[k]);
        *e = block_sum(B.n, [&](size_t i) {
            R kinetic = 0.5 * B.mass[i] * squaredNorm(B.vel[i]);
            R potential = dot(gv, Vec<R, D>::Zero() - B.pos[i]);
            return (double)(kinetic + B.mass[i] * potential);
        });
        return SPHX_OK;
    });
}
sphx_status sphx_upper_front_in_axis(sphx_body *body, int axis, double *front)
{
    if (!front)
        return fail(SPHX_E_ARG, "null argument");
    return with_body(body, [&](auto &B) -> sphx_status {
        BODY_TYPES(B);
        if (axis < 0 || axis >= D)
            return fail(SPHX_E_ARG, "axis out of range");
        *front = (double)block_max<R>(B.n, [&](size_t i) { return B.pos[i][axis]; });
        return SPHX_OK;
    });
}
// Interpolation<Contact<Real>>::interact, interpolation_dynamics.hpp:46-60
sphx_status sphx_interpolate_contact(sphx_relation *rel, int field, double *out, size_t n_obs)
{
    if (!rel || rel->inner || !rel->built || !out)
        return fail(SPHX_E_ARG, "interpolation needs an updated contact relation");
    return with_body(rel->source, [&](auto &O) -> sphx_status {
        BODY_TYPES(O);
        Body<R, D> &T = as<R, D>(rel->target);
        if (n_obs != O.n)
            return fail(SPHX_E_ARG, "observer count mismatch");
        const std::vector<R> *data = nullptr;
        NeighborKernel<R, D> nk(*O.kernel, O.h, T.h);
        if (field == SPHX_FIELD_VELOCITY)
        {
            // Interpolation<Contact<Vecd>>: weighted sum of the vector / (total weight + TinyReal)
            for (UInt i = 0; i < O.n; ++i)
            {
                Vec<R, D> q = Vec<R, D>::Zero();
                R w(0);
                for (UInt n = rel->particle_offset[i]; n != rel->particle_offset[i + 1]; ++n)
                {
                    UInt j = rel->neighbor_index[n];
                    R weight_j = nk.W(O.pos[i] - T.pos[j]) * T.Vol[j];
                    q += weight_j * T.vel[j];
                    w += weight_j;
                }
                Vec<R, D> v = q / (w + TinyReal<R>());
                for (int k = 0; k < D; ++k)
                    out[i * D + k] = (double)v[k];
            }
            return SPHX_OK;
        }
        if (field == SPHX_FIELD_PRESSURE)
            data = &T.p;
        else if (field == SPHX_FIELD_DENSITY)
            data = &T.rho;
        else if (field == SPHX_FIELD_COMPRESSION)
            data = &T.C;
        else
            return fail(SPHX_E_ARG, "unsupported interpolation field");
        for (UInt i = 0; i < O.n; ++i)
        {
            R q(0), w(0);
            for (UInt n = rel->particle_offset[i]; n != rel->particle_offset[i + 1]; ++n)
            {
                UInt j = rel->neighbor_index[n];
                R weight_j = nk.W(O.pos[i] - T.pos[j]) * T.Vol[j];
                q += weight_j * (*data)[j];
                w += weight_j;
            }
            O.interp[i] = q / (w + TinyReal<R>());
            out[i] = (double)O.interp[i];
        }
        return SPHX_OK;
    });
}

// the CK 3D dam-break loop body, dambreak.cpp:188-222, with the sort placed
// as the 3D case does (after the position update, before the CLL update)
sphx_status sphx_advection_step(sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                                sphx_cell_linked_list *cll, const sphx_advection_config *cfg,
                                double *time, uint32_t *steps)
{
    if (!cfg || !time || !relations_ok(inner, walls, n_wall) || !cll || cll != inner->cll)
        return fail(SPHX_E_ARG, "advection step arguments inconsistent");
    sphx_body *fluid = inner->source;
    sphx_status s;
    if ((s = sphx_compression_summation(inner, walls, n_wall)))
        return s;
    if ((s = sphx_density_regularization(fluid, cfg->flow_type)))
        return s;
    if ((s = sphx_advection_step_setup(fluid)))
        return s;
    double adv_dt = 0;
    if ((s = sphx_advection_time_step(fluid, cfg->u_ref, cfg->advection_cfl, &adv_dt)))
        return s;
    // dambreak.cpp:192-193
    if (cfg->surface_indication && (s = sphx_free_surface_indication(inner, walls, n_wall)))
        return s;
    const bool corrected = cfg->correction == SPHX_CORRECTION_LINEAR;
    if (corrected && (s = sphx_linear_correction_matrix(inner, walls, n_wall, cfg->correction_alpha)))
        return s;
    double relaxation_time = 0.0;
    uint32_t count = 0;
    while (relaxation_time < adv_dt)
    {
        double dt = 0;
        if ((s = sphx_acoustic_time_step(fluid, cfg->acoustic_cfl, &dt)))
            return s;
        if ((s = first_half(inner, walls, n_wall, cfg->riemann, dt, corrected)))
            return s;
        if ((s = second_half(inner, walls, n_wall, cfg->riemann, dt, corrected)))
            return s;
        relaxation_time += dt;
        *time += dt;
        ++count;
    }
    if ((s = sphx_update_particle_position(fluid)))
        return s;
    if (cfg->sort_now && (s = sphx_particle_sort(fluid, cll)))
        return s;
    if ((s = sphx_update_cell_linked_list(cll)))
        return s;
    if ((s = sphx_update_relation(inner)))
        return s;
    for (int k = 0; k < n_wall; ++k)
        if ((s = sphx_update_relation(walls[k])))
            return s;
    if (steps)
        *steps = count;
    return SPHX_OK;
}

// batched scenes on the CPU: each scene's advection step in turn, which is
// what the device batch must reproduce scene by scene
struct sphx_batch
{
    struct Scene
    {
        sphx_relation *inner;
        std::vector<sphx_relation *> walls;
        sphx_cell_linked_list *cll;
        sphx_advection_config cfg;
        bool failed = false;
    };
    std::vector<Scene> scenes;
    uint64_t steps = 0;
};

sphx_status sphx_batch_create(int, sphx_batch **out)
{
    if (!out)
        return fail(SPHX_E_ARG, "null out");
    *out = new sphx_batch();
    return SPHX_OK;
}
sphx_status sphx_batch_destroy(sphx_batch *b)
{
    delete b;
    return SPHX_OK;
}
sphx_status sphx_batch_set_options(sphx_batch *b, const sphx_batch_options *o)
{
    return b && o ? SPHX_OK : fail(SPHX_E_ARG, "null argument");
}
sphx_status sphx_batch_add_scene(sphx_batch *b, sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                                 sphx_cell_linked_list *cll, const sphx_advection_config *cfg)
{
    if (!b || !cfg || !relations_ok(inner, walls, n_wall) || !cll || cll != inner->cll)
        return fail(SPHX_E_ARG, "batched scene arguments inconsistent");
    sphx_batch::Scene sc{inner, std::vector<sphx_relation *>(walls, walls + n_wall), cll, *cfg};
    b->scenes.push_back(sc);
    return SPHX_OK;
}
sphx_status sphx_batch_remove_scene(sphx_batch *b, uint32_t index)
{
    if (!b || index >= b->scenes.size())
        return fail(SPHX_E_ARG, "no such scene in the batch");
    b->scenes.erase(b->scenes.begin() + index);
    return SPHX_OK;
}
sphx_status sphx_batch_advection_step(sphx_batch *b, double *time, uint32_t *steps, const int *sort_now, int *status)
{
    if (!b || !time)
        return fail(SPHX_E_ARG, "null argument");
    sphx_status result = SPHX_OK;
    for (size_t i = 0; i < b->scenes.size(); ++i)
    {
        sphx_batch::Scene &sc = b->scenes[i];
        sphx_advection_config cfg = sc.cfg;
        cfg.sort_now = sort_now ? sort_now[i] : 0;
        uint32_t n = 0;
        sphx_status s = sc.failed ? SPHX_E_NONFINITE
                                  : sphx_advection_step(sc.inner, sc.walls.data(), (int)sc.walls.size(), sc.cll, &cfg,
                                                        &time[i], &n);
        if (s && s != SPHX_E_NONFINITE)
            return s;
        if (s)
        {
            sc.failed = true;
            result = SPHX_E_NONFINITE;
        }
        if (steps)
            steps[i] = n;
        if (status)
            status[i] = s;
    }
    b->steps++;
    return result;
}
sphx_status sphx_batch_tuning(const sphx_batch *b, char *json, size_t capacity, size_t *length)
{
    if (!b)
        return fail(SPHX_E_ARG, "null batch");
    const std::string j = "{\"scenes\": " + std::to_string(b->scenes.size()) + ", \"steps\": " +
                          std::to_string(b->steps) + ", \"serial\": true}";
    if (length)
        *length = j.size();
    if (json && capacity > 0)
    {
        const size_t n = std::min(capacity - 1, j.size());
        std::memcpy(json, j.data(), n);
        json[n] = '\0';
    }
    return SPHX_OK;
}

// instrumentation entry points: the oracle records nothing
sphx_status sphx_profile_enable(sphx_context *ctx, int)
{
    return ctx ? SPHX_OK : fail(SPHX_E_ARG, "null context");
}
sphx_status sphx_profile_read(sphx_context *ctx, const char *stage, double *ms, uint64_t *launches)
{
    if (!ctx || !stage)
        return fail(SPHX_E_ARG, "null argument");
    if (ms)
        *ms = 0.0;
    if (launches)
        *launches = 0;
    return SPHX_OK;
}

// oracle-only: thread count used by parallel_for (the CPU baseline reports it)
int sphx_oracle_set_threads(int n)
{
    if (n > 0)
        thread_count() = n;
    return thread_count();
}

} // extern "C"

extern "C" {
// oracle-only: the exclusive scan of algorithm_primitive.h:244-279, for its
// known-answer test (tests/unit_tests_src/.../test_exclusive_scan.cpp:8-25)
uint32_t sphx_oracle_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n)
{
    return orc::exclusive_scan(in, out, n);
}
}
