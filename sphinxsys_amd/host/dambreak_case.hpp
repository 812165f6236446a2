// dambreak_case.hpp -- the SPHinXsys dam-break cases on the C++ CK mirror.
//
// Geometry of tests/2d_examples/test_2d_dambreak/Dambreak.cpp:12-47 and
// tests/tests_sycl/2d_examples/test_2d_dambreak_sycl/dambreak_sycl.cpp:12-62
// (2D), and of the 3D dam break (tank 3.2 x 1.0 x 1.5, column 1.2 x 1.0 x 0.6,
// 3 wall layers, gravity along -z):
//  - GeometricShapeBox contain / closest point (geometric_element.cpp:19-59,
//    transform_geometry.h:53-64), ComplexShape add/subtract
//    (base_geometry.cpp:84-137);
//  - ParticleGenerator<BaseParticles, Lattice> over the system bounds expanded
//    by 4 dp (particle_generator_lattice_3d.cpp:12-25, sph_system.cpp:39);
//  - Shape::findNormalDirection for the wall normals (base_geometry.cpp:45-58);
//  - particles emitted in the Morton order of their cell-linked-list cell.
// Scene and loop: dambreak.cpp:111-222 (CK 3D case), with the plain wall
// Riemann solver and no kernel correction.
// All arithmetic is IEEE double in the same operation order as the Python
// case generator, so both produce bit-identical particle sets.
#pragma once

#include "sphx_ck.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <vector>

namespace sphx
{
namespace dambreak
{
struct Box
{
    std::vector<double> t, h; // centre, half size
    bool contains(const double *p) const
    {
        for (size_t k = 0; k < t.size(); ++k)
            if (!(std::fabs(p[k] - t[k]) <= h[k]))
                return false;
        return true;
    }
    // Simbody Geo::Box closest point, as GeometricBox::findClosestPoint
    void closest_point(const double *p, double *out) const
    {
        const size_t d = t.size();
        double c[3];
        bool inside = true;
        for (size_t k = 0; k < d; ++k)
        {
            c[k] = p[k] - t[k];
            if (c[k] < -h[k])
            {
                c[k] = -h[k];
                inside = false;
            }
            else if (c[k] > h[k])
            {
                c[k] = h[k];
                inside = false;
            }
        }
        if (inside)
        {
            size_t which = 0;
            double mind = h[0] - std::fabs(c[0]);
            for (size_t k = 1; k < d; ++k)
            {
                const double dk = h[k] - std::fabs(c[k]);
                if (dk < mind)
                {
                    mind = dk;
                    which = k;
                }
            }
            c[which] = c[which] < 0 ? -h[which] : h[which];
        }
        for (size_t k = 0; k < d; ++k)
            out[k] = t[k] + c[k];
    }
};

struct ComplexShape
{
    std::vector<std::pair<Box, bool>> parts; // (shape, add?)
    ComplexShape &add(const Box &b)
    {
        parts.push_back({b, true});
        return *this;
    }
    ComplexShape &subtract(const Box &b)
    {
        parts.push_back({b, false});
        return *this;
    }
    bool contains(const double *p) const
    {
        bool exist = false;
        for (const auto &pr : parts)
        {
            const bool in = pr.first.contains(p);
            exist = pr.second ? (exist || in) : (exist && !in);
        }
        return exist;
    }
    void closest_point(const double *p, size_t d, double *out) const
    {
        double best = std::numeric_limits<double>::max();
        for (const auto &pr : parts)
        {
            double f[3];
            pr.first.closest_point(p, f);
            double s = 0.0;
            for (size_t k = 0; k < d; ++k)
                s = s + (p[k] - f[k]) * (p[k] - f[k]);
            const double dist = std::sqrt(s);
            if (dist <= best)
            {
                best = dist;
                for (size_t k = 0; k < d; ++k)
                    out[k] = f[k];
            }
        }
    }
};

// cell centres of a dp lattice over [lo, up) inside the shape, first axis slowest
template <class Shape>
std::vector<double> lattice_positions(const std::vector<double> &lo, const std::vector<double> &up, double dp,
                                      const Shape &shape)
{
    const size_t d = lo.size();
    std::vector<std::vector<double>> axes(d);
    for (size_t k = 0; k < d; ++k)
    {
        const long n = (long)std::ceil((up[k] + 0.0 - lo[k]) / dp);
        for (long i = 0; i < n; ++i)
            axes[k].push_back((lo[k] + (double)i * dp) + 0.5 * dp);
    }
    std::vector<double> out;
    double p[3] = {0.0, 0.0, 0.0};
    for (double x : axes[0])
        for (double y : axes[1])
        {
            if (d == 2)
            {
                p[0] = x;
                p[1] = y;
                if (shape.contains(p))
                    out.insert(out.end(), p, p + 2);
                continue;
            }
            for (double z : axes[2])
            {
                p[0] = x;
                p[1] = y;
                p[2] = z;
                if (shape.contains(p))
                    out.insert(out.end(), p, p + 3);
            }
        }
    return out;
}

// Morton order of the cell-linked-list cells (gs, lower bound - 2 gs), stable
inline std::vector<double> morton_ordered(const std::vector<double> &pos, size_t d, const std::vector<double> &lo,
                                          double gs)
{
    const size_t n = pos.size() / d;
    std::vector<uint64_t> key(n, 0);
    for (size_t i = 0; i < n; ++i)
    {
        for (size_t a = 0; a < d; ++a)
        {
            const double lower = lo[a] - 2.0 * gs;
            const int64_t c = std::max<int64_t>((int64_t)std::floor((pos[i * d + a] - lower) / gs), 0);
            for (int bit = 0; bit < 10; ++bit)
                key[i] |= (((uint64_t)c >> bit) & 1ull) << (bit * 3 + a);
        }
    }
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return key[x] < key[y]; });
    std::vector<double> out(pos.size());
    for (size_t i = 0; i < n; ++i)
        for (size_t a = 0; a < d; ++a)
            out[i * d + a] = pos[idx[i] * d + a];
    return out;
}

inline std::vector<double> normals(const ComplexShape &shape, const std::vector<double> &pos, size_t d)
{
    const size_t n = pos.size() / d;
    std::vector<double> out(pos.size());
    for (size_t i = 0; i < n; ++i)
    {
        const double *p = &pos[i * d];
        double c[3];
        shape.closest_point(p, d, c);
        double disp[3], z = 0.0;
        for (size_t k = 0; k < d; ++k)
        {
            disp[k] = c[k] - p[k];
            z = z + disp[k] * disp[k];
        }
        if (!(z > 0.0))
            throw ck::Error("normal requested on the shape surface");
        const double s = std::sqrt(z);
        const bool in = shape.contains(p);
        for (size_t k = 0; k < d; ++k)
            out[i * d + k] = in ? disp[k] / s : -1.0 * (disp[k] / s);
    }
    return out;
}

struct Geometry
{
    int dim = 2;
    double dp = 0.0;
    std::vector<double> fluid, wall, wall_normal, observer;
    std::vector<double> lower, upper; // system bounds
    double rho0 = 1.0, c0 = 1.0, u_ref = 1.0;
    std::vector<double> gravity;
};

// 2D dam break, plain WCSPH configuration (Dambreak.cpp:12-47)
inline Geometry dambreak_2d(double dp = 0.025, double DL = 5.366, double DH = 5.366, double LL = 2.0,
                            double LH = 1.0, int wall_layers = 4, double rho0 = 1.0, double g = 1.0)
{
    Geometry G;
    G.dim = 2;
    G.dp = dp;
    const double BW = dp * wall_layers;
    G.u_ref = 2.0 * std::sqrt(g * LH);
    G.c0 = 10.0 * G.u_ref;
    G.rho0 = rho0;
    Box water{{0.5 * LL, 0.5 * LH}, {0.5 * LL, 0.5 * LH}};
    const double oh0 = 0.5 * DL + BW, oh1 = 0.5 * DH + BW;
    ComplexShape wall;
    wall.add(Box{{-BW + oh0, -BW + oh1}, {oh0, oh1}}).subtract(Box{{0.5 * DL, 0.5 * DH}, {0.5 * DL, 0.5 * DH}});
    G.lower = {-BW - dp * 4, -BW - dp * 4};
    G.upper = {(DL + BW) + dp * 4, (DH + BW) + dp * 4};
    const double gs = 2.0 * 1.3 * dp;
    G.fluid = morton_ordered(lattice_positions(G.lower, G.upper, dp, water), 2, G.lower, gs);
    G.wall = morton_ordered(lattice_positions(G.lower, G.upper, dp, wall), 2, G.lower, gs);
    G.wall_normal = normals(wall, G.wall, 2);
    G.observer = {DL, 0.2};
    G.gravity = {0.0, -g};
    return G;
}

// 3D dam break: x length, y width, z height (gravity along -z)
inline Geometry dambreak_3d(double dp = 0.00448, double DL = 3.2, double DW = 1.0, double DH = 1.5,
                            double LL = 1.2, double LW = 1.0, double LH = 0.6, int wall_layers = 3,
                            double rho0 = 1000.0, double g = 9.81)
{
    Geometry G;
    G.dim = 3;
    G.dp = dp;
    const double BW = dp * wall_layers;
    G.u_ref = 2.0 * std::sqrt(g * LH);
    G.c0 = 10.0 * G.u_ref;
    G.rho0 = rho0;
    std::vector<double> hw = {0.5 * LL, 0.5 * LW, 0.5 * LH};
    Box water{hw, hw};
    std::vector<double> hin = {0.5 * DL, 0.5 * DW, 0.5 * DH};
    std::vector<double> hout = {0.5 * DL + BW, 0.5 * DW + BW, 0.5 * DH + BW};
    ComplexShape wall;
    wall.add(Box{hin, hout}).subtract(Box{hin, hin});
    G.lower = {-BW - dp * 4, -BW - dp * 4, -BW - dp * 4};
    G.upper = {(DL + BW) + dp * 4, (DW + BW) + dp * 4, (DH + BW) + dp * 4};
    const double gs = 2.0 * 1.3 * dp;
    G.fluid = morton_ordered(lattice_positions(G.lower, G.upper, dp, water), 3, G.lower, gs);
    G.wall = morton_ordered(lattice_positions(G.lower, G.upper, dp, wall), 3, G.lower, gs);
    G.wall_normal = normals(wall, G.wall, 3);
    G.observer = {3.2, 0.5, 0.1};
    G.gravity = {0.0, 0.0, -g};
    return G;
}

// Bodies, relations and dynamics of one dam-break scene (dambreak.cpp:111-154).
template <class ExecutionPolicy = ck::execution::ParallelHipPolicy>
class Scene
{
  public:
    using P = ExecutionPolicy;
    Scene(ck::SPHSystem &sys, const Geometry &G, ck::Precision precision)
        : G_(G),
          water_block(sys, props(G, SPHX_BODY_FLUID, precision), G.fluid.size() / G.dim),
          wall_boundary(sys, props(G, SPHX_BODY_SOLID, precision), G.wall.size() / G.dim),
          fluid_observer(sys, props(G, SPHX_BODY_OBSERVER, precision), G.observer.size() / G.dim),
          water_cll((init_bodies(), water_block), G.lower, G.upper), wall_cll(wall_boundary, G.lower, G.upper),
          water_block_inner(water_block, water_cll), water_wall_contact(water_block, wall_boundary, wall_cll),
          fluid_observer_contact(fluid_observer, water_block, water_cll),
          water_cell_linked_list(water_cll), wall_cell_linked_list(wall_cll),
          water_block_update_complex_relation(water_block_inner, water_wall_contact),
          fluid_observer_contact_relation(fluid_observer_contact), particle_sort(water_block, water_cll),
          constant_gravity(water_block, G.gravity),
          fluid_density_summation(water_block_inner, water_wall_contact),
          fluid_density_regularization(water_block), water_advection_step_setup(water_block),
          water_update_particle_position(water_block),
          fluid_acoustic_step_1st_half(water_block_inner, water_wall_contact),
          fluid_acoustic_step_2nd_half(water_block_inner, water_wall_contact),
          fluid_advection_time_step(water_block, G.u_ref), fluid_acoustic_time_step(water_block),
          total_energy(water_block, G.gravity), probe(fluid_observer_contact, "Pressure"),
          advection(water_block_inner, {&water_wall_contact}, water_cll, G.u_ref)
    {
    }

    // preparation, dambreak.cpp:148-154
    void initialize()
    {
        constant_gravity.exec();
        water_cell_linked_list.exec();
        wall_cell_linked_list.exec();
        water_block_update_complex_relation.exec();
        fluid_observer_contact_relation.exec();
    }

    // one advection step through the single-call driver; ParticleSortCK when the
    // incremented iteration count is a multiple of 100 (dambreak.cpp:210-216)
    uint32_t advection_step()
    {
        const bool sort = (iterations + 1) % 100 == 0;
        const uint32_t n = advection.exec(physical_time, sort);
        ++iterations;
        acoustic_steps += n;
        return n;
    }

    // the same step written out with one CK dynamics object per stage
    uint32_t advection_step_by_dynamics()
    {
        fluid_density_summation.exec();
        fluid_density_regularization.exec();
        water_advection_step_setup.exec();
        const double advection_dt = fluid_advection_time_step.exec();
        double relaxation_time = 0.0;
        uint32_t n = 0;
        while (relaxation_time < advection_dt)
        {
            const double dt = fluid_acoustic_time_step.exec();
            fluid_acoustic_step_1st_half.exec(dt);
            fluid_acoustic_step_2nd_half.exec(dt);
            relaxation_time += dt;
            physical_time += dt;
            ++n;
        }
        water_update_particle_position.exec();
        ++iterations;
        if (iterations % 100 == 0 && iterations != 1)
            particle_sort.exec();
        water_cell_linked_list.exec();
        water_block_update_complex_relation.exec();
        acoustic_steps += n;
        return n;
    }

    // probe pressure and total mechanical energy
    std::pair<double, double> observe()
    {
        fluid_observer_contact_relation.exec();
        const double p = probe.interact()[0];
        return {p, total_energy.exec()};
    }

  private:
    static ck::SPHBody::Properties props(const Geometry &G, int kind, ck::Precision precision)
    {
        ck::SPHBody::Properties p;
        p.dim = G.dim;
        p.kind = kind;
        p.precision = precision;
        p.spacing_ref = G.dp;
        p.rho0 = G.rho0;
        p.c0 = G.c0;
        return p;
    }
    int init_bodies()
    {
        const double vol = std::pow(G_.dp, G_.dim);
        water_block.write("Position", G_.fluid);
        water_block.write("VolumetricMeasure", std::vector<double>(water_block.size(), vol));
        wall_boundary.write("Position", G_.wall);
        wall_boundary.write("VolumetricMeasure", std::vector<double>(wall_boundary.size(), vol));
        wall_boundary.write("NormalDirection", G_.wall_normal);
        fluid_observer.write("Position", G_.observer);
        fluid_observer.write("VolumetricMeasure", std::vector<double>(fluid_observer.size(), vol));
        return 0;
    }
    const Geometry &G_;

  public:
    ck::RealBody water_block, wall_boundary, fluid_observer;
    ck::CellLinkedList water_cll, wall_cll;
    ck::Inner water_block_inner;
    ck::Contact water_wall_contact, fluid_observer_contact;
    ck::UpdateCellLinkedList<P, ck::RealBody> water_cell_linked_list, wall_cell_linked_list;
    ck::UpdateRelation<P, ck::Inner, ck::Contact> water_block_update_complex_relation;
    ck::UpdateRelation<P, ck::Contact> fluid_observer_contact_relation;
    ck::ParticleSortCK<P> particle_sort;
    ck::StateDynamics<P, ck::GravityForceCK> constant_gravity;
    ck::InteractionDynamicsCK<P, ck::fluid_dynamics::CompressionSummation> fluid_density_summation;
    ck::StateDynamics<P, ck::fluid_dynamics::DensityRegularization<ck::fluid_dynamics::FreeSurface>>
        fluid_density_regularization;
    ck::StateDynamics<P, ck::fluid_dynamics::AdvectionStepSetup> water_advection_step_setup;
    ck::StateDynamics<P, ck::fluid_dynamics::UpdateParticlePosition> water_update_particle_position;
    ck::InteractionDynamicsCK<P, ck::fluid_dynamics::AcousticStep1stHalfWithWallRiemannCK>
        fluid_acoustic_step_1st_half;
    ck::InteractionDynamicsCK<P, ck::fluid_dynamics::AcousticStep2ndHalfWithWallRiemannCK>
        fluid_acoustic_step_2nd_half;
    ck::ReduceDynamicsCK<P, ck::fluid_dynamics::AdvectionTimeStepCK> fluid_advection_time_step;
    ck::ReduceDynamicsCK<P, ck::fluid_dynamics::AcousticTimeStepCK> fluid_acoustic_time_step;
    ck::ReduceDynamicsCK<P, ck::TotalMechanicalEnergyCK> total_energy;
    ck::InterpolationContact probe;
    ck::AdvectionStep advection;
    double physical_time = 0.0;
    uint32_t iterations = 0, acoustic_steps = 0;
};

} // namespace dambreak
} // namespace sphx
