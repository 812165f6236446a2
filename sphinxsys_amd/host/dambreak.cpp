// dambreak.cpp -- the dam-break case on the MI355X through the C++ CK mirror
// (execution::par_hip).  Counterpart of the reference case programs
// tests/tests_sycl/2d_examples/test_2d_dambreak_sycl/dambreak_sycl.cpp and
// tests/tests_sycl/3d_examples/test_3d_dambreak_sycl/dambreak.cpp.
//
//   dambreak [--dim 2|3] [--dp DP] [--precision fp64|mixed] [--steps N]
//            [--t-end T] [--observe DT] [--by-dynamics]
//
// Prints one line per observation (physical time, probe pressure, total
// mechanical energy) and a summary line with the particle-steps per second
// of the stepping loop.
#include "dambreak_case.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

using namespace sphx;

int main(int argc, char **argv)
{
    int dim = 2, steps = -1;
    double dp = 0.0, t_end = -1.0, observe = -1.0;
    bool by_dynamics = false;
    ck::Precision precision = ck::Precision::FP64;
    for (int a = 1; a < argc; ++a)
    {
        const std::string k = argv[a];
        auto next = [&]() -> std::string {
            if (a + 1 >= argc)
                throw ck::Error("missing value after " + k);
            return argv[++a];
        };
        if (k == "--dim")
            dim = std::stoi(next());
        else if (k == "--dp")
            dp = std::stod(next());
        else if (k == "--precision")
            precision = next() == "mixed" ? ck::Precision::Mixed : ck::Precision::FP64;
        else if (k == "--steps")
            steps = std::stoi(next());
        else if (k == "--t-end")
            t_end = std::stod(next());
        else if (k == "--observe")
            observe = std::stod(next());
        else if (k == "--by-dynamics")
            by_dynamics = true;
        else
        {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
    }
    try
    {
        const dambreak::Geometry G = dim == 2 ? (dp > 0 ? dambreak::dambreak_2d(dp) : dambreak::dambreak_2d())
                                              : (dp > 0 ? dambreak::dambreak_3d(dp) : dambreak::dambreak_3d());
        if (steps < 0 && t_end < 0)
            t_end = dim == 2 ? 20.0 : 1.5; // end times of the reference cases
        if (observe <= 0)
            observe = dim == 2 ? 0.1 : 0.01;
        ck::SPHSystem system(ck::execution::par_hip);
        dambreak::Scene<> scene(system, G, precision);
        scene.initialize();
        std::printf("# %s, %zu fluid, %zu wall particles\n", system.api().backend().c_str(), scene.water_block.size(),
                    scene.wall_boundary.size());
        double next_obs = 0.0;
        auto obs = [&]() {
            const auto pe = scene.observe();
            std::printf("%.9g %.12g %.12g\n", scene.physical_time, pe.first, pe.second);
        };
        obs();
        next_obs += observe;
        const auto t0 = std::chrono::steady_clock::now();
        for (int it = 0; steps < 0 || it < steps; ++it)
        {
            if (steps < 0 && scene.physical_time >= t_end)
                break;
            if (by_dynamics)
                scene.advection_step_by_dynamics();
            else
                scene.advection_step();
            if (scene.physical_time >= next_obs)
            {
                obs();
                next_obs += observe;
            }
        }
        system.synchronize();
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("# advection steps %u, acoustic steps %u, %.3f s, %.4g particle-steps/s\n", scene.iterations,
                    scene.acoustic_steps, wall, (double)scene.water_block.size() * scene.acoustic_steps / wall);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "dambreak: %s\n", e.what());
        return 1;
    }
    return 0;
}
