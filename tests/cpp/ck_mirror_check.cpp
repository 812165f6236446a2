// ck_mirror_check.cpp -- drives the dam-break scene of the C++ CK mirror on
// the library given on the command line and dumps its state, so that the
// tests can compare it with the Python mirror on the same library.
//
//   ck_mirror_check LIBRARY DIM PRECISION(0|1) STEPS MODE(driver|dynamics) OUT_PREFIX
//
// Writes OUT_PREFIX.geometry.bin (fluid, wall, normal positions as doubles),
// OUT_PREFIX.state.bin (Position, Velocity, Compression of the fluid after
// STEPS advection steps) and prints "time probe energy acoustic_steps".
#include "../../sphinxsys_amd/host/dambreak_case.hpp"

#include <cstdio>
#include <string>

using namespace sphx;

static void dump(const std::string &path, std::initializer_list<const std::vector<double> *> parts)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        throw ck::Error("cannot write " + path);
    for (const auto *v : parts)
        std::fwrite(v->data(), sizeof(double), v->size(), f);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc != 7)
    {
        std::fprintf(stderr, "usage: %s LIBRARY DIM PRECISION STEPS MODE OUT_PREFIX\n", argv[0]);
        return 2;
    }
    try
    {
        const std::string lib = argv[1], mode = argv[5], out = argv[6];
        const int dim = std::stoi(argv[2]), steps = std::stoi(argv[4]);
        const auto precision = std::stoi(argv[3]) == 1 ? ck::Precision::Mixed : ck::Precision::FP64;
        const dambreak::Geometry G = dim == 2 ? dambreak::dambreak_2d() : dambreak::dambreak_3d(0.04);
        dump(out + ".geometry.bin", {&G.fluid, &G.wall, &G.wall_normal});
        ck::SPHSystem system(ck::Api::load(lib));
        dambreak::Scene<> scene(system, G, precision);
        scene.initialize();
        for (int it = 0; it < steps; ++it)
            mode == "dynamics" ? scene.advection_step_by_dynamics() : scene.advection_step();
        const auto pe = scene.observe();
        const auto pos = scene.water_block.read("Position"), vel = scene.water_block.read("Velocity"),
                   C = scene.water_block.read("Compression");
        dump(out + ".state.bin", {&pos, &vel, &C});
        std::printf("%.17g %.17g %.17g %u\n", scene.physical_time, pe.first, pe.second, scene.acoustic_steps);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "ck_mirror_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
