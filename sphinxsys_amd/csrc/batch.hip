// batch.hip -- many independent scenes advanced together on one gfx950 device.
//
// A batch keeps a scene table on the device (per scene: particle arrays,
// relation slices, parameters, time-step slots) and launches every stage of
// the advection step once over the blocks of all scenes; a block finds its
// scene and its place in that scene's grid through a block map.  The step is
// captured in a HIP graph and replayed, so a step of the whole batch is one
// graph launch plus one small read-back of the scenes' loop states.  Scenes
// the graph could not finish (more sub-steps than captured, neighbour lists
// that outgrew their width, a particle sort due) are finished through the
// single-scene path, which continues from the same device state.
//
// The autotuner varies launch shapes that do not change any result (the
// tiled neighbour build's waves per workgroup, which scenes run the acoustic
// passes from LDS-staged neighbour records) over the first steps and keeps
// the fastest by measured graph time per particle sub-step.
#include "sphx_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace sphx;

namespace
{
struct Shape
{
    int tile_waves = 1;
    uint32_t stage_min = 0; // scenes with at least this many fluid particles are staged
    bool operator==(const Shape &o) const { return tile_waves == o.tile_waves && stage_min == o.stage_min; }
};

constexpr uint32_t kMapLimit = 1u << 16; // scenes, relations and blocks per scene in a block map entry
} // namespace

struct sphx_batch
{
    int device = 0;
    hipStream_t stream = nullptr;
    int D = 0, P = 0, riemann = 0;
    sphx_batch_options opt{};
    struct Scene
    {
        sphx_relation *inner = nullptr;
        sphx_relation *walls[4] = {nullptr, nullptr, nullptr, nullptr};
        int n_wall = 0;
        sphx_cell_linked_list *cll = nullptr;
        sphx_advection_config cfg{};
        uint32_t last = 0;     // sub-steps of the previous step
        bool failed = false;   // diverged: stays without advancing
        bool staged = false;   // passes from LDS-staged records under the current shape
        uint32_t rel0 = 0;     // index of its inner relation in the relation table
    };
    std::vector<Scene> scenes;
    uint32_t n_rel = 0;
    bool tables = false;      // device tables and maps built for the current scene set
    bool tables_dirty = false;// host tables changed since the last upload
    Shape shape;              // shape the tables were filled for
    // device buffers
    std::vector<void *> allocations;
    void *scene_tab = nullptr, *cll_tab = nullptr, *build_tab = nullptr, *stage_tab = nullptr;
    uint32_t *map256 = nullptr, *map_cells = nullptr, *map_tiles = nullptr, *map_chunks[3] = {nullptr, nullptr, nullptr};
    uint32_t *rel_scene = nullptr, *cursor = nullptr, *tile_sums = nullptr, *tail_on = nullptr, *gate = nullptr;
    double *time_d = nullptr;
    unsigned long long *readback = nullptr; // S * STEP_SLOTS step states, then R needed counts
    uint32_t blocks256 = 0, blocks_cells = 0, blocks_tiles = 0, blocks_chunks[3] = {0, 0, 0};
    // pinned host mirrors
    double *time_h = nullptr;
    uint32_t *tail_on_h = nullptr;
    unsigned long long *readback_h = nullptr;
    std::vector<unsigned char> scene_tab_h, build_tab_h, stage_tab_h;
    // graph of one step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_substeps = 0;
    Shape graph_shape;
    bool graph_valid = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // tuning
    std::vector<Shape> candidates;
    std::vector<double> cost;   // ms per million particle sub-steps, per candidate
    size_t next_candidate = 0;
    bool tuned = false;
    // counters
    uint64_t steps = 0, graph_launches = 0, captures = 0, direct_steps = 0;
    uint64_t host_substep_scenes = 0, host_tail_scenes = 0, regrown_relations = 0;
};

namespace
{
sphx_status batch_error(sphx_status code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int wave_slot(int w) { return w == 1 ? 0 : (w == 2 ? 1 : 2); }

void free_tables(sphx_batch *b)
{
    for (void *p : b->allocations)
        dev_free(b->stream, p);
    b->allocations.clear();
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
    b->tables = false;
}

template <class T>
sphx_status alloc_dev(sphx_batch *b, T **p, size_t count)
{
    void *q = nullptr;
    sphx_status s = dev_alloc(b->stream, &q, std::max<size_t>(count, 1) * sizeof(T));
    if (s)
        return s;
    b->allocations.push_back(q);
    *p = (T *)q;
    return SPHX_OK;
}

// the per-scene records of the scene, build and staging tables for the current shape
void fill_records(sphx_batch *b)
{
    const size_t sa = batch_scene_args_bytes(b->P), ba = batch_build_args_bytes(), st = batch_stage_args_bytes();
    unsigned long long *needed = b->readback + (size_t)b->scenes.size() * STEP_SLOTS;
    for (size_t s = 0; s < b->scenes.size(); ++s)
    {
        sphx_batch::Scene &sc = b->scenes[s];
        sc.staged = sc.inner->staged && sc.inner->source->n >= b->shape.stage_min;
        batch_scene_args(b->scene_tab_h.data() + s * sa, sc.inner, sc.walls, sc.n_wall, &sc.cfg, b->time_d + s,
                         sc.staged);
        batch_stage_args(b->stage_tab_h.data() + s * st, sc.staged ? sc.inner : nullptr);
        batch_build_args(b->build_tab_h.data() + (size_t)sc.rel0 * ba, sc.inner, needed + sc.rel0);
        for (int k = 0; k < sc.n_wall; ++k)
            batch_build_args(b->build_tab_h.data() + (size_t)(sc.rel0 + 1 + k) * ba, sc.walls[k],
                             needed + sc.rel0 + 1 + k);
    }
    b->tables_dirty = true;
}

// scenes staged under the current shape whose staging sets are stale (their
// lists were last rebuilt without them) get them rebuilt on their own stream
sphx_status refresh_staging(sphx_batch *b)
{
    for (auto &sc : b->scenes)
        if (sc.staged && sc.inner->stage_stale)
        {
            sphx_status s = relation_stage(sc.inner);
            if (s)
                return s;
            SPHX_HIP(hipStreamSynchronize(sc.inner->source->ctx->stream));
            sc.inner->stage_stale = false;
        }
    return SPHX_OK;
}

sphx_status upload_records(sphx_batch *b)
{
    if (!b->tables_dirty)
        return SPHX_OK;
    sphx_status s = refresh_staging(b);
    if (s)
        return s;
    SPHX_HIP(hipMemcpyAsync(b->scene_tab, b->scene_tab_h.data(), b->scene_tab_h.size(), hipMemcpyHostToDevice,
                            b->stream));
    SPHX_HIP(hipMemcpyAsync(b->build_tab, b->build_tab_h.data(), b->build_tab_h.size(), hipMemcpyHostToDevice,
                            b->stream));
    SPHX_HIP(hipMemcpyAsync(b->stage_tab, b->stage_tab_h.data(), b->stage_tab_h.size(), hipMemcpyHostToDevice,
                            b->stream));
    // the host vectors are pageable: these copies complete before returning
    SPHX_HIP(hipStreamSynchronize(b->stream));
    b->tables_dirty = false;
    return SPHX_OK;
}

// block maps, tables and work arrays for the current scene set
sphx_status build_tables(sphx_batch *b)
{
    free_tables(b);
    const size_t S = b->scenes.size();
    std::vector<uint32_t> m256, mcells, mtiles, mchunks[3], rscene;
    std::vector<unsigned char> cll_h(S * batch_cll_args_bytes());
    size_t cells_total = 0, tiles_total = 0;
    for (size_t s = 0; s < S; ++s)
    {
        const sphx_batch::Scene &sc = b->scenes[s];
        const uint32_t n = sc.inner->source->n;
        const uint64_t nc = sc.cll->ncells;
        const uint64_t b256 = (n + 255u) / 256u, bcells = (nc + 1 + 255u) / 256u;
        const uint32_t tiles = batch_cll_tiles(sc.cll);
        if (b256 > kMapLimit || bcells > kMapLimit || tiles > kMapLimit)
            return batch_error(SPHX_E_CAPACITY, "a batched scene exceeds 65536 blocks of particles or cells");
        for (uint32_t k = 0; k < b256; ++k)
            m256.push_back((uint32_t)(s << 16) | k);
        for (uint32_t k = 0; k < bcells; ++k)
            mcells.push_back((uint32_t)(s << 16) | k);
        for (uint32_t k = 0; k < tiles; ++k)
            mtiles.push_back((uint32_t)(s << 16) | k);
        cells_total += nc;
        tiles_total += tiles;
    }
    for (size_t s = 0; s < S; ++s)
    {
        const sphx_batch::Scene &sc = b->scenes[s];
        const uint32_t nchunks = (sc.inner->source->n + 63u) / 64u;
        for (int r = 0; r <= sc.n_wall; ++r)
        {
            const uint32_t rel = sc.rel0 + (uint32_t)r;
            rscene.push_back((uint32_t)s);
            for (int w = 0; w < 3; ++w)
            {
                const uint32_t waves = 1u << w, groups = (nchunks + waves - 1) / waves;
                if (groups > kMapLimit)
                    return batch_error(SPHX_E_CAPACITY, "a batched relation exceeds 65536 workgroups");
                for (uint32_t k = 0; k < groups; ++k)
                    mchunks[w].push_back((rel << 16) | k);
            }
        }
    }
    sphx_status st;
    const size_t R = b->n_rel;
    if ((st = alloc_dev(b, (unsigned char **)&b->scene_tab, S * batch_scene_args_bytes(b->P))) ||
        (st = alloc_dev(b, (unsigned char **)&b->cll_tab, S * batch_cll_args_bytes())) ||
        (st = alloc_dev(b, (unsigned char **)&b->build_tab, R * batch_build_args_bytes())) ||
        (st = alloc_dev(b, (unsigned char **)&b->stage_tab, S * batch_stage_args_bytes())) ||
        (st = alloc_dev(b, &b->map256, m256.size())) || (st = alloc_dev(b, &b->map_cells, mcells.size())) ||
        (st = alloc_dev(b, &b->map_tiles, mtiles.size())) || (st = alloc_dev(b, &b->map_chunks[0], mchunks[0].size())) ||
        (st = alloc_dev(b, &b->map_chunks[1], mchunks[1].size())) ||
        (st = alloc_dev(b, &b->map_chunks[2], mchunks[2].size())) || (st = alloc_dev(b, &b->rel_scene, R)) ||
        (st = alloc_dev(b, &b->cursor, cells_total)) || (st = alloc_dev(b, &b->tile_sums, tiles_total)) ||
        (st = alloc_dev(b, &b->tail_on, S)) || (st = alloc_dev(b, &b->gate, S)) || (st = alloc_dev(b, &b->time_d, S)) ||
        (st = alloc_dev(b, &b->readback, S * STEP_SLOTS + R)))
        return st;
    // cell-list records: per-scene slices of the cursor and tile-sum arrays
    size_t c0 = 0, t0 = 0;
    for (size_t s = 0; s < S; ++s)
    {
        sphx_cell_linked_list *cll = b->scenes[s].cll;
        batch_cll_args(cll_h.data() + s * batch_cll_args_bytes(), cll, b->cursor + c0, b->tile_sums + t0);
        c0 += cll->ncells;
        t0 += batch_cll_tiles(cll);
    }
    auto up = [&](void *dst, const void *src, size_t bytes) -> sphx_status {
        SPHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream));
        return SPHX_OK;
    };
    if ((st = up(b->map256, m256.data(), m256.size() * 4)) || (st = up(b->map_cells, mcells.data(), mcells.size() * 4)) ||
        (st = up(b->map_tiles, mtiles.data(), mtiles.size() * 4)) ||
        (st = up(b->map_chunks[0], mchunks[0].data(), mchunks[0].size() * 4)) ||
        (st = up(b->map_chunks[1], mchunks[1].data(), mchunks[1].size() * 4)) ||
        (st = up(b->map_chunks[2], mchunks[2].data(), mchunks[2].size() * 4)) ||
        (st = up(b->rel_scene, rscene.data(), rscene.size() * 4)) || (st = up(b->cll_tab, cll_h.data(), cll_h.size())))
        return st;
    SPHX_HIP(hipStreamSynchronize(b->stream));
    b->blocks256 = (uint32_t)m256.size();
    b->blocks_cells = (uint32_t)mcells.size();
    b->blocks_tiles = (uint32_t)mtiles.size();
    for (int w = 0; w < 3; ++w)
        b->blocks_chunks[w] = (uint32_t)mchunks[w].size();
    b->scene_tab_h.assign(S * batch_scene_args_bytes(b->P), 0);
    b->build_tab_h.assign(R * batch_build_args_bytes(), 0);
    b->stage_tab_h.assign(S * batch_stage_args_bytes(), 0);
    fill_records(b);
    b->tables = true;
    return upload_records(b);
}

BatchFluid fluid_launch(sphx_batch *b)
{
    BatchFluid f;
    f.D = b->D;
    f.P = b->P;
    f.riemann = b->riemann;
    f.stream = b->stream;
    f.tab = b->scene_tab;
    f.n_scenes = (uint32_t)b->scenes.size();
    f.map = b->map256;
    f.blocks = b->blocks256;
    f.stage_cap = 0;
    for (const auto &sc : b->scenes)
        if (sc.staged)
            f.stage_cap = std::max(f.stage_cap, sc.inner->stage_cap);
    f.tail_on = b->tail_on;
    f.gate = b->gate;
    f.state = b->readback;
    return f;
}

BatchConfig config_launch(sphx_batch *b)
{
    BatchConfig c;
    c.D = b->D;
    c.stream = b->stream;
    c.gate = b->gate;
    c.cll = b->cll_tab;
    c.n_scenes = (uint32_t)b->scenes.size();
    c.map_particles = b->map256;
    c.blocks_particles = b->blocks256;
    c.map_cells = b->map_cells;
    c.blocks_cells = b->blocks_cells;
    c.map_tiles = b->map_tiles;
    c.blocks_tiles = b->blocks_tiles;
    c.build = b->build_tab;
    c.n_rel = b->n_rel;
    const int w = wave_slot(b->shape.tile_waves);
    c.map_chunks = b->map_chunks[w];
    c.blocks_chunks = b->blocks_chunks[w];
    c.rel_scene = b->rel_scene;
    c.needed = b->readback + b->scenes.size() * STEP_SLOTS;
    c.tile_waves = b->shape.tile_waves;
    bool any = false;
    for (const auto &sc : b->scenes)
        any |= sc.staged;
    c.stage = any ? b->stage_tab : nullptr;
    return c;
}

// the launches of the step's tail: position update, cell lists, relations
sphx_status enqueue_tail(sphx_batch *b)
{
    sphx_status s;
    const BatchFluid f = fluid_launch(b);
    const BatchConfig c = config_launch(b);
    if ((s = batch_step_gate(f)) || (s = batch_position(f)) || (s = batch_cell_linked_list(c)) ||
        (s = batch_relations(c)))
        return s;
    return SPHX_OK;
}

// the whole step with `substeps` captured sub-steps
sphx_status enqueue_step(sphx_batch *b, uint32_t substeps)
{
    const BatchFluid f = fluid_launch(b);
    sphx_status s = batch_step_head(f);
    for (uint32_t k = 0; k < substeps && !s; ++k)
        s = batch_substep(f);
    return s ? s : enqueue_tail(b);
}

sphx_status read_back(sphx_batch *b, bool with_needed)
{
    const size_t words = b->scenes.size() * STEP_SLOTS + (with_needed ? b->n_rel : 0);
    SPHX_HIP(hipMemcpyAsync(b->readback_h, b->readback, words * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            b->stream));
    SPHX_HIP(hipStreamSynchronize(b->stream));
    return SPHX_OK;
}

double state_double(const unsigned long long *st, int slot)
{
    double d;
    std::memcpy(&d, st + slot, sizeof(d));
    return d;
}

bool loop_over(const unsigned long long *st)
{
    return st[STEP_ERROR] != 0 || !(state_double(st, STEP_RELAXATION) < state_double(st, STEP_ADV_DT));
}

// direct launches with the loop checked on the host after each batch of sub-steps
sphx_status run_direct(sphx_batch *b)
{
    sphx_status s;
    const BatchFluid f = fluid_launch(b);
    if ((s = batch_step_head(f)))
        return s;
    uint32_t batch = 1;
    for (const auto &sc : b->scenes)
        batch = std::max(batch, sc.last);
    while (true)
    {
        for (uint32_t k = 0; k < batch; ++k)
            if ((s = batch_substep(f)))
                return s;
        if ((s = batch_step_gate(f)) || (s = read_back(b, false)))
            return s;
        bool all = true;
        for (size_t i = 0; i < b->scenes.size() && all; ++i)
            all = loop_over(b->readback_h + i * STEP_SLOTS);
        if (all)
            break;
        batch = 1;
    }
    if ((s = enqueue_tail(b)) || (s = read_back(b, true)))
        return s;
    b->direct_steps++;
    return SPHX_OK;
}

// the captured step, captured again when its sub-step count or shape changed;
// *ms is the graph's device time
sphx_status run_graph(sphx_batch *b, uint32_t substeps, float *ms)
{
    sphx_status s;
    if (!b->graph_valid || b->graph_substeps != substeps || !(b->graph_shape == b->shape))
    {
        if (b->exec)
            (void)hipGraphExecDestroy(b->exec);
        if (b->graph)
            (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->graph_valid = false;
        SPHX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        s = enqueue_step(b, substeps);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (s)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return s;
        }
        if (e != hipSuccess)
            return hip_fail(e, "hipStreamEndCapture");
        b->graph = g;
        SPHX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_substeps = substeps;
        b->graph_shape = b->shape;
        b->graph_valid = true;
        b->captures++;
    }
    SPHX_HIP(hipEventRecord(b->ev0, b->stream));
    SPHX_HIP(hipGraphLaunch(b->exec, b->stream));
    SPHX_HIP(hipEventRecord(b->ev1, b->stream));
    if ((s = read_back(b, true)))
        return s;
    SPHX_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
    b->graph_launches++;
    return SPHX_OK;
}

// host-side bookkeeping of a scene whose tail ran on the device, and the
// relations that outgrew their slice width rebuilt through the single-scene path
sphx_status finish_device_tail(sphx_batch *b, sphx_batch::Scene &sc, bool *regrown)
{
    sphx_body *body = sc.inner->source;
    body->state_version++;
    body->config_version++;
    sphx_cell_linked_list *c = sc.cll;
    c->built = true;
    ++c->build_count;
    c->built_version = body->config_version;
    const unsigned long long *needed = b->readback_h + b->scenes.size() * STEP_SLOTS;
    for (int r = 0; r <= sc.n_wall; ++r)
    {
        sphx_relation *rel = r == 0 ? sc.inner : sc.walls[r - 1];
        if (needed[sc.rel0 + r] > rel->kmax)
        {
            sphx_status s = sphx_update_relation(rel); // grows the width (or takes the scan build) and rebuilds
            if (s)
                return s;
            *regrown = true;
            b->regrown_relations++;
            continue;
        }
        rel->csr_valid = false;
        rel->built = true;
        rel->builds++;
        rel->built_version = rel->source->config_version;
        rel->target_version = rel->target->config_version;
        if (rel->staged)
            rel->stage_stale = !(r == 0 && sc.staged);
    }
    return SPHX_OK;
}

void tuning_candidates(sphx_batch *b)
{
    b->candidates.clear();
    b->cost.clear();
    b->next_candidate = 0;
    b->tuned = !b->opt.autotune;
    if (!b->opt.autotune)
        return;
    // which scenes are staged (all, from 4096, 16384 or 65536 fluid particles;
    // larger scenes always are, as in the single-scene path, whose 3D passes
    // depend on it), then the tiled build's waves per workgroup for the best
    for (uint32_t smin : {0u, 4096u, 16384u, 65536u})
        b->candidates.push_back(Shape{1, smin});
}

std::string shape_json(const Shape &s)
{
    char buf[96];
    std::snprintf(buf, sizeof(buf), "{\"tile_waves\": %d, \"stage_min\": %u}", s.tile_waves, s.stage_min);
    return buf;
}
} // namespace

extern "C" {

sphx_status sphx_batch_create(int device, sphx_batch **out)
{
    if (!out)
        return batch_error(SPHX_E_ARG, "null out");
    int count = 0;
    SPHX_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return batch_error(SPHX_E_ARG, "device index out of range");
    SPHX_HIP(hipSetDevice(device));
    std::unique_ptr<sphx_batch> b(new sphx_batch());
    b->device = device;
    b->opt.use_graph = 1;
    b->opt.autotune = 1;
    b->opt.tile_waves = 1;
    b->opt.stage_min = 0;
    b->opt.substep_margin = 1;
    struct Undo
    {
        sphx_batch *b;
        ~Undo()
        {
            if (!b)
                return;
            if (b->ev0)
                (void)hipEventDestroy(b->ev0);
            if (b->ev1)
                (void)hipEventDestroy(b->ev1);
            if (b->time_h)
                (void)hipHostFree(b->time_h);
            if (b->tail_on_h)
                (void)hipHostFree(b->tail_on_h);
            if (b->stream)
                (void)hipStreamDestroy(b->stream);
        }
    } undo{b.get()};
    SPHX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    SPHX_HIP(hipEventCreate(&b->ev0));
    SPHX_HIP(hipEventCreate(&b->ev1));
    SPHX_HIP(hipHostMalloc((void **)&b->time_h, 8, hipHostMallocDefault));
    SPHX_HIP(hipHostMalloc((void **)&b->tail_on_h, 8, hipHostMallocDefault));
    undo.b = nullptr;
    tuning_candidates(b.get());
    *out = b.release();
    return SPHX_OK;
}

sphx_status sphx_batch_destroy(sphx_batch *b)
{
    if (!b)
        return SPHX_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_tables(b);
    (void)hipStreamSynchronize(b->stream);
    (void)hipEventDestroy(b->ev0);
    (void)hipEventDestroy(b->ev1);
    (void)hipHostFree(b->time_h);
    (void)hipHostFree(b->tail_on_h);
    if (b->readback_h)
        (void)hipHostFree(b->readback_h);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return SPHX_OK;
}

sphx_status sphx_batch_set_options(sphx_batch *b, const sphx_batch_options *o)
{
    if (!b || !o)
        return batch_error(SPHX_E_ARG, "null argument");
    if (o->tile_waves != 1 && o->tile_waves != 2 && o->tile_waves != 4)
        return batch_error(SPHX_E_ARG, "tile_waves must be 1, 2 or 4");
    b->opt = *o;
    b->shape = Shape{o->tile_waves, o->stage_min};
    tuning_candidates(b);
    if (b->tables)
        fill_records(b);
    return SPHX_OK;
}

sphx_status sphx_batch_add_scene(sphx_batch *b, sphx_relation *inner, sphx_relation *const *walls, int n_wall,
                                 sphx_cell_linked_list *cll, const sphx_advection_config *cfg)
{
    if (!b || !cfg)
        return batch_error(SPHX_E_ARG, "null argument");
    sphx_status s = check_fluid_relations(inner, walls, n_wall, true);
    if (s)
        return s;
    if (!cll || cll != inner->cll)
        return batch_error(SPHX_E_ARG, "a batched scene needs its fluid cell linked list");
    if (cfg->correction != SPHX_CORRECTION_NONE || cfg->surface_indication)
        return batch_error(SPHX_E_ARG, "batched scenes run without kernel correction and surface indication");
    sphx_body *body = inner->source;
    if (body->ctx->device != b->device)
        return batch_error(SPHX_E_ARG, "the scene lives on another device than the batch");
    if (body->n == 0)
        return batch_error(SPHX_E_ARG, "a batched scene needs fluid particles");
    if (b->scenes.empty())
    {
        b->D = body->D;
        b->P = body->P;
        b->riemann = cfg->riemann;
    }
    else if (b->D != body->D || b->P != body->P || (b->riemann == SPHX_RIEMANN_NONE) != (cfg->riemann == SPHX_RIEMANN_NONE))
        return batch_error(SPHX_E_ARG, "batched scenes share dimension, precision and Riemann solver use");
    if (b->scenes.size() + 1 >= kMapLimit || b->n_rel + 1 + (uint32_t)n_wall >= kMapLimit)
        return batch_error(SPHX_E_CAPACITY, "too many scenes in one batch");
    sphx_batch::Scene sc;
    sc.inner = inner;
    sc.n_wall = n_wall;
    for (int k = 0; k < n_wall; ++k)
        sc.walls[k] = walls[k];
    sc.cll = cll;
    sc.cfg = *cfg;
    sc.cfg.sort_now = 0;
    sc.last = body->ctx->last_substeps;
    sc.rel0 = b->n_rel;
    b->n_rel += 1 + (uint32_t)n_wall;
    b->scenes.push_back(sc);
    // the tables and the pinned mirrors are sized on the next step
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_tables(b);
    return SPHX_OK;
}

sphx_status sphx_batch_remove_scene(sphx_batch *b, uint32_t index)
{
    if (!b || index >= b->scenes.size())
        return batch_error(SPHX_E_ARG, "no such scene in the batch");
    SPHX_HIP(hipSetDevice(b->device));
    SPHX_HIP(hipStreamSynchronize(b->stream));
    b->scenes.erase(b->scenes.begin() + index);
    b->n_rel = 0;
    for (auto &sc : b->scenes)
    {
        sc.rel0 = b->n_rel;
        b->n_rel += 1 + (uint32_t)sc.n_wall;
    }
    free_tables(b);
    return SPHX_OK;
}

sphx_status sphx_batch_advection_step(sphx_batch *b, double *time, uint32_t *steps, const int *sort_now, int *status)
{
    if (!b || !time)
        return batch_error(SPHX_E_ARG, "null argument");
    if (b->scenes.empty())
        return SPHX_OK;
    SPHX_HIP(hipSetDevice(b->device));
    const size_t S = b->scenes.size();
    sphx_status s;
    for (const auto &sc : b->scenes)
        if ((s = check_fluid_relations(sc.inner, sc.walls, sc.n_wall, true)))
            return s;
    if (!b->tables)
    {
        // pinned mirrors of the per-step inputs and of the read-back
        (void)hipHostFree(b->time_h);
        (void)hipHostFree(b->tail_on_h);
        if (b->readback_h)
            (void)hipHostFree(b->readback_h);
        b->time_h = nullptr;
        b->tail_on_h = nullptr;
        b->readback_h = nullptr;
        SPHX_HIP(hipHostMalloc((void **)&b->time_h, S * sizeof(double), hipHostMallocDefault));
        SPHX_HIP(hipHostMalloc((void **)&b->tail_on_h, S * sizeof(uint32_t), hipHostMallocDefault));
        SPHX_HIP(hipHostMalloc((void **)&b->readback_h, (S * STEP_SLOTS + b->n_rel) * sizeof(unsigned long long),
                               hipHostMallocDefault));
        if (!b->opt.autotune)
            b->shape = Shape{b->opt.tile_waves, b->opt.stage_min};
        else if (!b->tuned && b->next_candidate < b->candidates.size())
            b->shape = b->candidates[b->next_candidate];
        // work the scenes' set-up left on their own streams
        for (auto &sc : b->scenes)
            SPHX_HIP(hipStreamSynchronize(sc.inner->source->ctx->stream));
        if ((s = build_tables(b)))
            return s;
    }
    // the tuning step's shape
    bool timing = false;
    if (b->opt.autotune && !b->tuned && b->steps > 0 && b->opt.use_graph)
    {
        if (b->next_candidate < b->candidates.size())
        {
            if (!(b->shape == b->candidates[b->next_candidate]))
            {
                b->shape = b->candidates[b->next_candidate];
                fill_records(b);
            }
            timing = true;
        }
    }
    if ((s = upload_records(b)))
        return s;
    for (size_t i = 0; i < S; ++i)
    {
        b->time_h[i] = time[i];
        b->tail_on_h[i] = (sort_now && sort_now[i]) || b->scenes[i].failed ? 0u : 1u;
    }
    SPHX_HIP(hipMemcpyAsync(b->time_d, b->time_h, S * sizeof(double), hipMemcpyHostToDevice, b->stream));
    SPHX_HIP(hipMemcpyAsync(b->tail_on, b->tail_on_h, S * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    // the first step (no sub-step counts yet) and the graph-less mode check the loop on the host
    uint32_t most = 0;
    for (const auto &sc : b->scenes)
        most = std::max(most, sc.failed ? 0u : sc.last);
    float ms = 0.0f;
    const bool graph = b->opt.use_graph && most > 0;
    if (graph)
        s = run_graph(b, most + b->opt.substep_margin, &ms);
    else
        s = run_direct(b);
    if (s)
        return s;
    // per scene: results, then what the host finishes
    sphx_status result = SPHX_OK;
    bool regrown = false;
    double particle_substeps = 0.0;
    for (size_t i = 0; i < S; ++i)
    {
        sphx_batch::Scene &sc = b->scenes[i];
        sphx_body *body = sc.inner->source;
        sphx_context *ctx = body->ctx;
        const unsigned long long *st = b->readback_h + i * STEP_SLOTS;
        int scene_status = SPHX_OK;
        if (st[STEP_ERROR] || sc.failed)
        {
            body->state_version++;
            sc.failed = true;
            scene_status = SPHX_E_NONFINITE;
        }
        else if (!loop_over(st))
        {
            // more sub-steps than captured: the single-scene loop goes on from the device state
            body->state_version++;
            std::memcpy(ctx->scalars_h + SLOT_STEP, st, STEP_SLOTS * sizeof(unsigned long long));
            if ((s = advection_substeps(sc.inner, sc.walls, sc.n_wall, &sc.cfg, 1u)))
            {
                if (s != SPHX_E_NONFINITE)
                    return s;
                sc.failed = true;
                scene_status = SPHX_E_NONFINITE;
            }
            else if ((s = advection_tail(sc.inner, sc.walls, sc.n_wall, sc.cll, sort_now && sort_now[i])))
                return s;
            b->host_substep_scenes++;
            b->host_tail_scenes++;
            regrown = true; // relations may have been regrown; the records are refreshed
            st = ctx->scalars_h + SLOT_STEP;
        }
        else if (!b->tail_on_h[i])
        {
            // the tail with the particle sort
            body->state_version++;
            if ((s = advection_tail(sc.inner, sc.walls, sc.n_wall, sc.cll, true)))
                return s;
            b->host_tail_scenes++;
            regrown = true;
        }
        else if ((s = finish_device_tail(b, sc, &regrown)))
            return s;
        if (status)
            status[i] = scene_status;
        if (scene_status != SPHX_OK)
        {
            result = SPHX_E_NONFINITE;
            if (steps)
                steps[i] = 0;
            continue;
        }
        const uint32_t count = (uint32_t)st[STEP_COUNT];
        time[i] = state_double(st, STEP_TIME);
        if (steps)
            steps[i] = count;
        sc.last = count;
        ctx->last_substeps = count;
        particle_substeps += (double)body->n * count;
    }
    // eager work on the scenes' own streams ends before the next batch step
    if (regrown)
    {
        for (auto &sc : b->scenes)
            SPHX_HIP(hipStreamSynchronize(sc.inner->source->ctx->stream));
        fill_records(b);
    }
    if (timing)
    {
        b->cost.push_back(particle_substeps > 0.0 ? 1e6 * ms / particle_substeps : 1e30);
        b->next_candidate++;
        if (b->next_candidate == 4 && b->candidates.size() == 4)
        {
            // best staging rule, then the build's waves per workgroup
            const size_t best = (size_t)(std::min_element(b->cost.begin(), b->cost.end()) - b->cost.begin());
            const uint32_t smin = b->candidates[best].stage_min;
            b->candidates.push_back(Shape{2, smin});
            b->candidates.push_back(Shape{4, smin});
        }
        if (b->next_candidate >= b->candidates.size())
        {
            const size_t best = (size_t)(std::min_element(b->cost.begin(), b->cost.end()) - b->cost.begin());
            b->tuned = true;
            if (!(b->shape == b->candidates[best]))
            {
                b->shape = b->candidates[best];
                fill_records(b);
            }
        }
    }
    if (s = upload_records(b); s)
        return s;
    b->steps++;
    return result;
}

sphx_status sphx_batch_tuning(const sphx_batch *b, char *json, size_t capacity, size_t *length)
{
    if (!b)
        return batch_error(SPHX_E_ARG, "null batch");
    std::string j = "{\"scenes\": " + std::to_string(b->scenes.size()) + ", \"autotune\": " +
                    std::to_string(b->opt.autotune) + ", \"tuned\": " + (b->tuned ? "true" : "false") +
                    ", \"candidates\": [";
    for (size_t k = 0; k < b->candidates.size(); ++k)
    {
        char cost[64] = "null";
        if (k < b->cost.size())
            std::snprintf(cost, sizeof(cost), "%.6g", b->cost[k]);
        j += (k ? ", " : "") + std::string("{\"shape\": ") + shape_json(b->candidates[k]) +
             ", \"ms_per_million_particle_substeps\": " + cost + "}";
    }
    j += "], \"shape\": " + shape_json(b->shape) + ", \"graph_substeps\": " + std::to_string(b->graph_substeps) +
         ", \"steps\": " + std::to_string(b->steps) + ", \"graph_launches\": " + std::to_string(b->graph_launches) +
         ", \"captures\": " + std::to_string(b->captures) + ", \"direct_steps\": " + std::to_string(b->direct_steps) +
         ", \"host_substep_scenes\": " + std::to_string(b->host_substep_scenes) +
         ", \"host_tail_scenes\": " + std::to_string(b->host_tail_scenes) +
         ", \"regrown_relations\": " + std::to_string(b->regrown_relations) + "}";
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

} // extern "C"
