// microbench.cpp -- neighbour search + density summation microbenchmark on the
// MI355X through the C++ CK mirror: N = m^3 particles on a lattice of spacing
// dp = 1/m filling the unit cube, x = (i + 1/2) dp per axis, each component
// jittered by U(-0.1 dp, 0.1 dp); h = 1.3 dp, cut-off 2h, cell linked list
// over the unit cube expanded by 4 dp.  The jitter comes from
// std::mt19937_64(seed) with u = (x >> 11) * 2^-53 (the generator's output
// sequence is fixed by the C++ standard, so every platform draws the same
// positions).  Stages timed with the library's HIP-event profile:
// UpdateCellLinkedList, ParticleSortCK (Morton sort + reorder), UpdateRelation
// (neighbour counts and lists) and CompressionSummation (the density sum).
//
//   microbench [--m 256] [--seed 20261202] [--precision fp64|mixed|both]
//              [--repeat 5] [--dump FILE]
//
// Prints one JSON object per precision.
#include "sphx_ck.hpp"

#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

using namespace sphx;

static std::vector<double> jittered_lattice(int m, uint64_t seed)
{
    const double dp = 1.0 / m;
    std::mt19937_64 rng(seed);
    std::vector<double> pos((size_t)m * m * m * 3);
    size_t k = 0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            for (int l = 0; l < m; ++l)
            {
                const int idx[3] = {i, j, l};
                for (int a = 0; a < 3; ++a)
                {
                    const double u = (double)(rng() >> 11) * 0x1.0p-53; // [0, 1)
                    pos[k++] = ((double)idx[a] + 0.5) * dp + (-0.1 * dp + 0.2 * dp * u);
                }
            }
    return pos;
}

static void run(const std::vector<double> &pos, int m, ck::Precision precision, int repeat)
{
    const double dp = 1.0 / m;
    const size_t n = pos.size() / 3;
    ck::SPHSystem system(ck::execution::par_hip);
    ck::SPHBody::Properties props;
    props.dim = 3;
    props.precision = precision;
    props.spacing_ref = dp;
    ck::RealBody body(system, props, n);
    body.write("Position", pos);
    body.write("VolumetricMeasure", std::vector<double>(n, dp * dp * dp));
    ck::CellLinkedList cll(body, {-4 * dp, -4 * dp, -4 * dp}, {1 + 4 * dp, 1 + 4 * dp, 1 + 4 * dp});
    ck::Inner inner(body, cll);
    using P = ck::execution::ParallelHipPolicy;
    ck::UpdateCellLinkedList<P> cell_linked_list(cll);
    ck::ParticleSortCK<P> particle_sort(body, cll);
    ck::UpdateRelation<P, ck::Inner> update_relation(inner);
    ck::InteractionDynamicsCK<P, ck::fluid_dynamics::CompressionSummation> density_sum(inner);
    // first pass from the lattice order: cell list, Morton sort, cell list again
    cell_linked_list.exec();
    particle_sort.exec();
    cell_linked_list.exec();
    update_relation.exec();
    density_sum.exec();
    system.synchronize();
    system.profile(true);
    for (int r = 0; r < repeat; ++r)
    {
        cell_linked_list.exec();
        particle_sort.exec();
        cell_linked_list.exec();
        update_relation.exec();
        density_sum.exec();
    }
    system.synchronize();
    auto ms = [&](const char *stage) { return system.profile_read(stage).first / repeat; };
    const double cll_ms = ms("cell_linked_list") / 2.0; // two builds per repetition
    const double sort_ms = ms("particle_sort");
    const double rel_ms = ms("relation_build") + ms("relation_scan") + ms("relation_stage");
    const double sum_ms = ms("compression_summation");
    system.profile(false);
    std::vector<uint32_t> size, offset, index;
    inner.read(size, offset, index);
    const auto S = body.read("CompressionSummation");
    double s_mean = 0.0;
    for (double v : S)
        s_mean += v;
    s_mean /= (double)n;
    std::printf("{\"precision\": \"%s\", \"particles\": %zu, \"cell_linked_list_ms\": %.4f, "
                "\"particle_sort_ms\": %.4f, \"relation_ms\": %.4f, \"density_sum_ms\": %.4f, "
                "\"neighbours_total\": %zu, \"neighbours_mean\": %.4f, \"summation_mean\": %.9f, "
                "\"particles_per_s_relation\": %.4g}\n",
                precision == ck::Precision::FP64 ? "fp64" : "fp32-mixed", n, cll_ms, sort_ms, rel_ms, sum_ms,
                index.size(), (double)index.size() / (double)n, s_mean, (double)n / (rel_ms * 1e-3));
    std::fflush(stdout);
}

int main(int argc, char **argv)
{
    int m = 256, repeat = 5;
    uint64_t seed = 20261202;
    std::string precision = "both", dump;
    for (int a = 1; a < argc; ++a)
    {
        const std::string k = argv[a];
        if (a + 1 >= argc)
        {
            std::fprintf(stderr, "missing value after %s\n", k.c_str());
            return 2;
        }
        const std::string v = argv[++a];
        if (k == "--m")
            m = std::stoi(v);
        else if (k == "--seed")
            seed = std::stoull(v);
        else if (k == "--precision")
            precision = v;
        else if (k == "--repeat")
            repeat = std::stoi(v);
        else if (k == "--dump")
            dump = v;
        else
        {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
    }
    try
    {
        const std::vector<double> pos = jittered_lattice(m, seed);
        if (!dump.empty())
        {
            FILE *f = std::fopen(dump.c_str(), "wb");
            if (!f)
                throw ck::Error("cannot write " + dump);
            std::fwrite(pos.data(), sizeof(double), pos.size(), f);
            std::fclose(f);
        }
        if (precision == "fp64" || precision == "both")
            run(pos, m, ck::Precision::FP64, repeat);
        if (precision == "mixed" || precision == "both")
            run(pos, m, ck::Precision::Mixed, repeat);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "microbench: %s\n", e.what());
        return 1;
    }
    return 0;
}
