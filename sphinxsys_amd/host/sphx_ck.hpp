// sphx_ck.hpp -- C++ mirror of the SPHinXsys computing-kernel (CK) dynamics
// classes used by the weakly-compressible fluid step, on top of the sphx C ABI
// (include/sphx.h).
//
// A SPHinXsys case file declares its dynamics as, for example,
//     UpdateCellLinkedList<MainExecutionPolicy, RealBody> water_cell_linked_list(water_block);
//     InteractionDynamicsCK<MainExecutionPolicy, AcousticStep1stHalfWithWallRiemannCK>
//         fluid_acoustic_step_1st_half(water_block_inner, water_wall_contact);
// (tests/tests_sycl/3d_examples/test_3d_dambreak_sycl/dambreak.cpp:111-136).
// Here the same declarations take execution::ParallelHipPolicy and the
// objects forward exec() to the device library.  Class names, constructor
// arguments and exec() follow the reference; errors throw sphx::ck::Error
// carrying sphx_last_error().
//
// The library is loaded at run time (dlopen) into a function table, so a
// program can drive any library that exports the sphx ABI.  execution::par_hip
// binds the HIP build, found through the SPHX_LIBRARY environment variable or
// the SPHX_DEFAULT_LIBRARY path compiled in; if it cannot be loaded the
// constructor throws (there is no CPU fallback).
#pragma once

#include "../../include/sphx.h"

#include <dlfcn.h>
#include <unistd.h>

#include <array>
#include <cstdlib>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sphx
{
namespace ck
{
class Error : public std::runtime_error
{
  public:
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

//-----------------------------------------------------------------------------
// function table of one library exporting the sphx ABI
//-----------------------------------------------------------------------------
#define SPHX_CK_FUNCTIONS(X)                                                                              \
    X(sphx_backend_name)                                                                                  \
    X(sphx_last_error)                                                                                    \
    X(sphx_context_create)                                                                                \
    X(sphx_context_destroy)                                                                               \
    X(sphx_context_synchronize)                                                                           \
    X(sphx_body_create)                                                                                   \
    X(sphx_body_destroy)                                                                                  \
    X(sphx_body_write)                                                                                    \
    X(sphx_body_read)                                                                                     \
    X(sphx_body_smoothing_length)                                                                         \
    X(sphx_cell_linked_list_create)                                                                       \
    X(sphx_cell_linked_list_destroy)                                                                      \
    X(sphx_cell_linked_list_mesh)                                                                         \
    X(sphx_update_cell_linked_list)                                                                       \
    X(sphx_cell_linked_list_read)                                                                         \
    X(sphx_relation_create_inner)                                                                         \
    X(sphx_relation_create_contact)                                                                       \
    X(sphx_relation_destroy)                                                                              \
    X(sphx_update_relation)                                                                               \
    X(sphx_relation_read)                                                                                 \
    X(sphx_particle_sort)                                                                                 \
    X(sphx_gravity_force)                                                                                 \
    X(sphx_density_regularization)                                                                        \
    X(sphx_advection_step_setup)                                                                          \
    X(sphx_update_particle_position)                                                                      \
    X(sphx_compression_summation)                                                                         \
    X(sphx_acoustic_step_1st_half)                                                                        \
    X(sphx_acoustic_step_2nd_half)                                                                        \
    X(sphx_acoustic_time_step)                                                                            \
    X(sphx_advection_time_step)                                                                           \
    X(sphx_total_mechanical_energy)                                                                       \
    X(sphx_upper_front_in_axis)                                                                           \
    X(sphx_interpolate_contact)                                                                           \
    X(sphx_advection_step)                                                                                \
    X(sphx_profile_enable)                                                                                \
    X(sphx_profile_read)                                                                                  \
    X(sphx_linear_correction_matrix)                                                                      \
    X(sphx_acoustic_step_1st_half_corrected)                                                              \
    X(sphx_acoustic_step_2nd_half_corrected)                                                              \
    X(sphx_free_surface_indication)                                                                       \
    X(sphx_viscous_force)                                                                                 \
    X(sphx_kernel_gradient_integral)                                                                      \
    X(sphx_transport_velocity_correction)

class Api
{
  public:
#define SPHX_CK_MEMBER(name) decltype(&::name) name = nullptr;
    SPHX_CK_FUNCTIONS(SPHX_CK_MEMBER)
#undef SPHX_CK_MEMBER

    // Loads `path` and resolves every entry point; throws on any failure.
    static std::shared_ptr<Api> load(const std::string &path)
    {
        void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h)
        {
            const char *e = dlerror();
            throw Error("cannot load " + path + ": " + (e ? e : "unknown error"));
        }
        std::shared_ptr<Api> api(new Api(h));
#define SPHX_CK_RESOLVE(name)                                                                            \
    api->name = reinterpret_cast<decltype(&::name)>(dlsym(h, #name));                                    \
    if (!api->name)                                                                                       \
        throw Error(path + " does not export " #name);
        SPHX_CK_FUNCTIONS(SPHX_CK_RESOLVE)
#undef SPHX_CK_RESOLVE
        return api;
    }
    ~Api()
    {
        if (handle_)
            dlclose(handle_);
    }
    Api(const Api &) = delete;
    Api &operator=(const Api &) = delete;

    void check(sphx_status s) const
    {
        if (s != SPHX_OK)
            throw Error(std::string("sphx status ") + std::to_string(s) + ": " + sphx_last_error());
    }
    std::string backend() const { return sphx_backend_name(); }

  private:
    explicit Api(void *h) : handle_(h) {}
    void *handle_ = nullptr;
};

//-----------------------------------------------------------------------------
// execution policies (execution_policy.h:53-83 shape)
//-----------------------------------------------------------------------------
namespace execution
{
class HipPolicy
{
};
template <typename...>
class DeviceExecution;
template <typename PolicyType>
class DeviceExecution<PolicyType> : public PolicyType
{
};
using ParallelHipPolicy = DeviceExecution<HipPolicy>;
inline constexpr auto par_hip = ParallelHipPolicy{};
} // namespace execution

// path of the HIP library: $SPHX_LIBRARY, else the compiled-in default, else
// libsphx.so next to the running executable
inline std::string hip_library_path()
{
    if (const char *e = std::getenv("SPHX_LIBRARY"))
        return e;
#ifdef SPHX_DEFAULT_LIBRARY
    return SPHX_DEFAULT_LIBRARY;
#else
    char buf[4096];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (n > 0)
    {
        std::string exe(buf, (size_t)n);
        const size_t slash = exe.rfind('/');
        if (slash != std::string::npos)
            return exe.substr(0, slash + 1) + "libsphx.so";
    }
    return "libsphx.so";
#endif
}

//-----------------------------------------------------------------------------
// system, bodies, cell linked lists, relations
//-----------------------------------------------------------------------------
enum class Precision
{
    FP64 = SPHX_FP64,
    Mixed = SPHX_MIXED
};

// the device of one SPHSystem: a context on one GPU
class SPHSystem
{
  public:
    explicit SPHSystem(execution::ParallelHipPolicy, int device = 0) : SPHSystem(load_hip(), device) {}
    SPHSystem(std::shared_ptr<Api> api, int device = 0) : api_(std::move(api))
    {
        api_->check(api_->sphx_context_create(device, &h_));
    }
    ~SPHSystem()
    {
        if (h_)
            api_->sphx_context_destroy(h_);
    }
    SPHSystem(const SPHSystem &) = delete;
    SPHSystem &operator=(const SPHSystem &) = delete;
    const Api &api() const { return *api_; }
    sphx_context *handle() const { return h_; }
    void synchronize() const { api_->check(api_->sphx_context_synchronize(h_)); }
    // HIP-event timing of the library's stages (sphx_profile_enable / _read)
    void profile(bool enable) const { api_->check(api_->sphx_profile_enable(h_, enable ? 1 : 0)); }
    std::pair<double, uint64_t> profile_read(const std::string &stage) const
    {
        double ms = 0.0;
        uint64_t n = 0;
        api_->check(api_->sphx_profile_read(h_, stage.c_str(), &ms, &n));
        return {ms, n};
    }

  private:
    static std::shared_ptr<Api> load_hip()
    {
        auto api = Api::load(hip_library_path());
        if (api->backend() != "hip-gfx950")
            throw Error("unexpected backend " + api->backend() + " at " + hip_library_path());
        return api;
    }
    std::shared_ptr<Api> api_;
    sphx_context *h_ = nullptr;
};

// registered-variable names of BaseParticles -> field ids
inline int field_id(const std::string &name)
{
    static const std::map<std::string, int> ids = {
        {"Position", SPHX_FIELD_POSITION},
        {"VolumetricMeasure", SPHX_FIELD_VOLUME},
        {"Mass", SPHX_FIELD_MASS},
        {"Density", SPHX_FIELD_DENSITY},
        {"Pressure", SPHX_FIELD_PRESSURE},
        {"Velocity", SPHX_FIELD_VELOCITY},
        {"Force", SPHX_FIELD_FORCE},
        {"ForcePrior", SPHX_FIELD_FORCE_PRIOR},
        {"Displacement", SPHX_FIELD_DISPLACEMENT},
        {"Compression", SPHX_FIELD_COMPRESSION},
        {"CompressionRate", SPHX_FIELD_COMPRESSION_RATE},
        {"VolumetricMeasureRef", SPHX_FIELD_VOLUME_REF},
        {"CompressionSummation", SPHX_FIELD_COMPRESSION_SUM},
        {"NormalDirection", SPHX_FIELD_NORMAL},
        {"WallVelocity", SPHX_FIELD_WALL_VELOCITY},
        {"WallAcceleration", SPHX_FIELD_WALL_ACCELERATION},
        {"OriginalID", SPHX_FIELD_ORIGINAL_ID},
        {"SortedID", SPHX_FIELD_SORTED_ID},
        {"SortSequence", SPHX_FIELD_SORT_SEQUENCE},
        {"Interpolated", SPHX_FIELD_INTERPOLATED},
        {"LinearCorrectionMatrix", SPHX_FIELD_CORRECTION_MATRIX},
        {"Indicator", SPHX_FIELD_INDICATOR},
        {"PreviousSurfaceIndicator", SPHX_FIELD_PREVIOUS_INDICATOR},
        {"PositionDivergence", SPHX_FIELD_POSITION_DIVERGENCE},
        {"ViscousForce", SPHX_FIELD_VISCOUS_FORCE},
        {"PreviousViscousForce", SPHX_FIELD_PREVIOUS_VISCOUS_FORCE},
        {"KernelGradientIntegral", SPHX_FIELD_KERNEL_GRADIENT_INTEGRAL},
    };
    auto it = ids.find(name);
    if (it == ids.end())
        throw Error("unknown particle variable " + name);
    return it->second;
}

inline bool is_vector_field(int id)
{
    return id == SPHX_FIELD_POSITION || id == SPHX_FIELD_VELOCITY || id == SPHX_FIELD_FORCE ||
           id == SPHX_FIELD_FORCE_PRIOR || id == SPHX_FIELD_DISPLACEMENT || id == SPHX_FIELD_NORMAL ||
           id == SPHX_FIELD_WALL_VELOCITY || id == SPHX_FIELD_WALL_ACCELERATION || id == SPHX_FIELD_VISCOUS_FORCE ||
           id == SPHX_FIELD_PREVIOUS_VISCOUS_FORCE || id == SPHX_FIELD_KERNEL_GRADIENT_INTEGRAL;
}
inline bool is_index_field(int id)
{
    return id == SPHX_FIELD_ORIGINAL_ID || id == SPHX_FIELD_SORTED_ID || id == SPHX_FIELD_SORT_SEQUENCE;
}
// int32 fields (read_int / write of std::vector<int32_t>)
inline bool is_int_field(int id) { return id == SPHX_FIELD_INDICATOR || id == SPHX_FIELD_PREVIOUS_INDICATOR; }

// a body with its particles (FluidBody / SolidBody / ObserverBody)
class SPHBody
{
  public:
    struct Properties
    {
        int dim = 3;
        int kind = SPHX_BODY_FLUID;
        Precision precision = Precision::FP64;
        double spacing_ref = 0.0;
        double h_spacing_ratio = 1.3; // SPHAdaptation default
        double rho0 = 1.0;
        double c0 = 1.0;
    };
    SPHBody(SPHSystem &system, const Properties &p, size_t n_particles) : system_(system), p_(p), n_(n_particles)
    {
        sphx_body_desc d{p.dim, (int)p.precision, p.kind, (uint32_t)n_particles, p.spacing_ref, p.h_spacing_ratio,
                         p.rho0, p.c0};
        api().check(api().sphx_body_create(system.handle(), &d, &h_));
    }
    ~SPHBody()
    {
        if (h_)
            api().sphx_body_destroy(h_);
    }
    SPHBody(const SPHBody &) = delete;
    SPHBody &operator=(const SPHBody &) = delete;

    const Api &api() const { return system_.api(); }
    sphx_body *handle() const { return h_; }
    SPHSystem &system() const { return system_; }
    int dim() const { return p_.dim; }
    size_t size() const { return n_; }
    double smoothing_length() const
    {
        double h = 0.0;
        api().check(api().sphx_body_smoothing_length(h_, &h));
        return h;
    }
    // real fields: n values, or n * dim AoS values for vectors
    void write(const std::string &variable, const std::vector<double> &v)
    {
        const int id = field_id(variable);
        expect(v.size(), id);
        api().check(api().sphx_body_write(h_, id, v.data(), n_));
    }
    void write(const std::string &variable, const std::vector<uint32_t> &v)
    {
        const int id = field_id(variable);
        expect(v.size(), id);
        api().check(api().sphx_body_write(h_, id, v.data(), n_));
    }
    void write(const std::string &variable, const std::vector<int32_t> &v)
    {
        const int id = field_id(variable);
        if (!is_int_field(id) || v.size() != n_)
            throw Error(variable + ": expected " + std::to_string(n_) + " int32 values of an int field");
        api().check(api().sphx_body_write(h_, id, v.data(), n_));
    }
    std::vector<double> read(const std::string &variable) const
    {
        const int id = field_id(variable);
        if (is_index_field(id) || is_int_field(id))
            throw Error(variable + " is an integer field: use read_index or read_int");
        std::vector<double> v(values_per_particle(id) * n_);
        api().check(api().sphx_body_read(h_, id, v.data(), n_));
        return v;
    }
    std::vector<int32_t> read_int(const std::string &variable) const
    {
        const int id = field_id(variable);
        if (!is_int_field(id))
            throw Error(variable + " is not an int field");
        std::vector<int32_t> v(n_);
        api().check(api().sphx_body_read(h_, id, v.data(), n_));
        return v;
    }
    std::vector<uint32_t> read_index(const std::string &variable) const
    {
        const int id = field_id(variable);
        if (!is_index_field(id))
            throw Error(variable + " is not an index field");
        std::vector<uint32_t> v(n_);
        api().check(api().sphx_body_read(h_, id, v.data(), n_));
        return v;
    }

  private:
    size_t values_per_particle(int id) const
    {
        if (id == SPHX_FIELD_CORRECTION_MATRIX)
            return (size_t)p_.dim * p_.dim;
        return is_vector_field(id) ? (size_t)p_.dim : 1;
    }
    void expect(size_t got, int id) const
    {
        const size_t want = values_per_particle(id) * n_;
        if (got != want)
            throw Error("field size " + std::to_string(got) + ", expected " + std::to_string(want));
    }
    SPHSystem &system_;
    Properties p_;
    size_t n_;
    sphx_body *h_ = nullptr;
};
using RealBody = SPHBody;

// CellLinkedList<SPHAdaptation> of a body over the system bounds, grid
// spacing = cut-off radius 2h, 2-cell buffer (cell_linked_list.cpp:190-197)
class CellLinkedList
{
  public:
    CellLinkedList(SPHBody &body, const std::vector<double> &lower, const std::vector<double> &upper,
                   double grid_spacing = 0.0, uint32_t buffer_width = 2)
        : body_(body)
    {
        if (grid_spacing <= 0.0)
            grid_spacing = 2.0 * body.smoothing_length();
        api().check(api().sphx_cell_linked_list_create(body.handle(), lower.data(), upper.data(), grid_spacing,
                                                        buffer_width, &h_));
    }
    ~CellLinkedList()
    {
        if (h_)
            api().sphx_cell_linked_list_destroy(h_);
    }
    CellLinkedList(const CellLinkedList &) = delete;
    CellLinkedList &operator=(const CellLinkedList &) = delete;
    const Api &api() const { return body_.api(); }
    sphx_cell_linked_list *handle() const { return h_; }
    SPHBody &body() const { return body_; }

  private:
    SPHBody &body_;
    sphx_cell_linked_list *h_ = nullptr;
};

class Relation
{
  public:
    ~Relation()
    {
        if (h_)
            src_.api().sphx_relation_destroy(h_);
    }
    Relation(const Relation &) = delete;
    Relation &operator=(const Relation &) = delete;
    sphx_relation *handle() const { return h_; }
    SPHBody &getSPHBody() const { return src_; }
    // NeighborSize, ParticleOffset, NeighborIndex (the reference's CSR layout)
    void read(std::vector<uint32_t> &size, std::vector<uint32_t> &offset, std::vector<uint32_t> &index) const
    {
        const Api &a = src_.api();
        size.resize(src_.size());
        offset.resize(src_.size() + 1);
        size_t total = 0;
        a.check(a.sphx_relation_read(h_, size.data(), offset.data(), nullptr, 0, &total));
        index.resize(total);
        a.check(a.sphx_relation_read(h_, size.data(), offset.data(), index.data(), index.size(), &total));
    }

  protected:
    explicit Relation(SPHBody &src) : src_(src) {}
    SPHBody &src_;
    sphx_relation *h_ = nullptr;
};

// Inner<> (relation_ck.h:89-124)
class Inner : public Relation
{
  public:
    Inner(SPHBody &body, CellLinkedList &cll) : Relation(body)
    {
        body.api().check(body.api().sphx_relation_create_inner(body.handle(), cll.handle(), &h_));
    }
};

// Contact<> (relation_ck.h:126-174), one contact body
class Contact : public Relation
{
  public:
    Contact(SPHBody &source, SPHBody &target, CellLinkedList &target_cll) : Relation(source)
    {
        source.api().check(
            source.api().sphx_relation_create_contact(source.handle(), target.handle(), target_cll.handle(), &h_));
    }
};

//-----------------------------------------------------------------------------
// dynamics: the CK class names, one exec() each
//-----------------------------------------------------------------------------
class BaseDynamics
{
  public:
    virtual ~BaseDynamics() = default;
};

template <class ExecutionPolicy, class DynamicsIdentifier = RealBody>
class UpdateCellLinkedList : public BaseDynamics
{
  public:
    explicit UpdateCellLinkedList(CellLinkedList &cll) : cll_(cll) {}
    void exec(double = 0.0) { cll_.api().check(cll_.api().sphx_update_cell_linked_list(cll_.handle())); }

  private:
    CellLinkedList &cll_;
};

template <class ExecutionPolicy, class... RelationTypes>
class UpdateRelation : public BaseDynamics
{
  public:
    explicit UpdateRelation(RelationTypes &...relations) : rel_{&relations...} {}
    void exec(double = 0.0)
    {
        for (Relation *r : rel_)
            r->getSPHBody().api().check(r->getSPHBody().api().sphx_update_relation(r->handle()));
    }

  private:
    std::vector<Relation *> rel_;
};

template <class ExecutionPolicy>
class ParticleSortCK : public BaseDynamics
{
  public:
    ParticleSortCK(SPHBody &body, CellLinkedList &cll) : body_(body), cll_(cll) {}
    void exec(double = 0.0) { body_.api().check(body_.api().sphx_particle_sort(body_.handle(), cll_.handle())); }

  private:
    SPHBody &body_;
    CellLinkedList &cll_;
};

// state dynamics of the fluid (StateDynamics<P, ...>)
class GravityForceCK
{
  public:
    GravityForceCK(SPHBody &body, const std::vector<double> &gravity) : body_(body), g_(gravity) {}
    void update() { body_.api().check(body_.api().sphx_gravity_force(body_.handle(), g_.data())); }

  private:
    SPHBody &body_;
    std::vector<double> g_;
};

namespace fluid_dynamics
{
struct FreeSurface
{
    static constexpr int flow_type = SPHX_FLOW_FREE_SURFACE;
};
struct Internal
{
    static constexpr int flow_type = SPHX_FLOW_INTERNAL;
};

template <class FlowType>
class DensityRegularization
{
  public:
    explicit DensityRegularization(SPHBody &body) : body_(body) {}
    void update()
    {
        body_.api().check(body_.api().sphx_density_regularization(body_.handle(), FlowType::flow_type));
    }

  private:
    SPHBody &body_;
};

class AdvectionStepSetup
{
  public:
    explicit AdvectionStepSetup(SPHBody &body) : body_(body) {}
    void update() { body_.api().check(body_.api().sphx_advection_step_setup(body_.handle())); }

  private:
    SPHBody &body_;
};

class UpdateParticlePosition
{
  public:
    explicit UpdateParticlePosition(SPHBody &body) : body_(body) {}
    void update() { body_.api().check(body_.api().sphx_update_particle_position(body_.handle())); }

  private:
    SPHBody &body_;
};

// interaction dynamics (InteractionDynamicsCK<P, ...>): inner + wall contacts
class FluidInteraction
{
  public:
    FluidInteraction(Inner &inner, std::initializer_list<Contact *> walls) : inner_(inner)
    {
        for (Contact *w : walls)
            walls_.push_back(w->handle());
    }

  protected:
    const Api &api() const { return inner_.getSPHBody().api(); }
    sphx_relation *const *walls() const { return walls_.empty() ? nullptr : walls_.data(); }
    Inner &inner_;
    std::vector<sphx_relation *> walls_;
};

class CompressionSummation : public FluidInteraction
{
  public:
    CompressionSummation(Inner &inner) : FluidInteraction(inner, {}) {}
    CompressionSummation(Inner &inner, Contact &wall) : FluidInteraction(inner, {&wall}) {}
    void interact()
    {
        api().check(api().sphx_compression_summation(inner_.handle(), walls(), (int)walls_.size()));
    }
};

// CORRECTION: SPHX_CORRECTION_NONE (NoKernelCorrectionCK) or SPHX_CORRECTION_LINEAR (LinearCorrectionCK)
template <int RIEMANN, int CORRECTION = SPHX_CORRECTION_NONE>
class AcousticStep1stHalfCK : public FluidInteraction
{
  public:
    AcousticStep1stHalfCK(Inner &inner) : FluidInteraction(inner, {}) {}
    AcousticStep1stHalfCK(Inner &inner, Contact &wall) : FluidInteraction(inner, {&wall}) {}
    void interact(double dt)
    {
        auto f = CORRECTION == SPHX_CORRECTION_LINEAR ? api().sphx_acoustic_step_1st_half_corrected
                                                      : api().sphx_acoustic_step_1st_half;
        api().check(f(inner_.handle(), walls(), (int)walls_.size(), RIEMANN, dt));
    }
};

template <int RIEMANN, int CORRECTION = SPHX_CORRECTION_NONE>
class AcousticStep2ndHalfCK : public FluidInteraction
{
  public:
    AcousticStep2ndHalfCK(Inner &inner) : FluidInteraction(inner, {}) {}
    AcousticStep2ndHalfCK(Inner &inner, Contact &wall) : FluidInteraction(inner, {&wall}) {}
    void interact(double dt)
    {
        auto f = CORRECTION == SPHX_CORRECTION_LINEAR ? api().sphx_acoustic_step_2nd_half_corrected
                                                      : api().sphx_acoustic_step_2nd_half;
        api().check(f(inner_.handle(), walls(), (int)walls_.size(), RIEMANN, dt));
    }
};

// acoustic_step_1st_half.h:196-201 / acoustic_step_2nd_half.h:183-191
using AcousticStep1stHalfWithWallRiemannCK = AcousticStep1stHalfCK<SPHX_RIEMANN_ACOUSTIC>;
using AcousticStep2ndHalfWithWallRiemannCK = AcousticStep2ndHalfCK<SPHX_RIEMANN_ACOUSTIC>;
using AcousticStep1stHalfWithWallNoRiemannCK = AcousticStep1stHalfCK<SPHX_RIEMANN_NONE>;
using AcousticStep2ndHalfWithWallNoRiemannCK = AcousticStep2ndHalfCK<SPHX_RIEMANN_NONE>;
using AcousticStep1stHalfWithWallRiemannCorrectionCK =
    AcousticStep1stHalfCK<SPHX_RIEMANN_ACOUSTIC, SPHX_CORRECTION_LINEAR>;
using AcousticStep2ndHalfWithWallRiemannCorrectionCK =
    AcousticStep2ndHalfCK<SPHX_RIEMANN_ACOUSTIC, SPHX_CORRECTION_LINEAR>;

// LinearCorrectionMatrix<Inner<WithUpdate>, Contact<>> (kernel_correction_ck.h:40-120), built as
// LinearCorrectionMatrixComplex(DynamicsArgs(inner, alpha), wall)
class LinearCorrectionMatrixComplex : public FluidInteraction
{
  public:
    LinearCorrectionMatrixComplex(Inner &inner, double alpha = 0.0) : FluidInteraction(inner, {}), alpha_(alpha) {}
    LinearCorrectionMatrixComplex(Inner &inner, double alpha, Contact &wall)
        : FluidInteraction(inner, {&wall}), alpha_(alpha)
    {
    }
    void interact()
    {
        api().check(api().sphx_linear_correction_matrix(inner_.handle(), walls(), (int)walls_.size(), alpha_));
    }

  private:
    double alpha_;
};

// ViscousForceCK<Inner<WithUpdate, Viscosity, K>, Contact<Wall, Viscosity, K>> (viscous_force.h:43-127)
// with the body's constant viscosity; ViscousForceWithWallCK is K = NoKernelCorrectionCK
template <int CORRECTION = SPHX_CORRECTION_NONE>
class ViscousForceCK : public FluidInteraction
{
  public:
    ViscousForceCK(Inner &inner, double mu) : FluidInteraction(inner, {}), mu_(mu) {}
    ViscousForceCK(Inner &inner, Contact &wall, double mu) : FluidInteraction(inner, {&wall}), mu_(mu) {}
    void interact()
    {
        api().check(api().sphx_viscous_force(inner_.handle(), walls(), (int)walls_.size(), mu_, CORRECTION));
    }

  private:
    double mu_;
};
using ViscousForceWithWallCK = ViscousForceCK<SPHX_CORRECTION_NONE>;

// KernelGradientIntegral<Inner<K>, Contact<Boundary, K>> (kernel_gradient_integral.h:89-93)
template <int CORRECTION>
class KernelGradientIntegralCK : public FluidInteraction
{
  public:
    KernelGradientIntegralCK(Inner &inner) : FluidInteraction(inner, {}) {}
    KernelGradientIntegralCK(Inner &inner, Contact &boundary) : FluidInteraction(inner, {&boundary}) {}
    void interact()
    {
        api().check(api().sphx_kernel_gradient_integral(inner_.handle(), walls(), (int)walls_.size(), CORRECTION));
    }
};
using KernelGradientIntegralComplex = KernelGradientIntegralCK<SPHX_CORRECTION_NONE>;
using KernelGradientIntegralCorrectedComplex = KernelGradientIntegralCK<SPHX_CORRECTION_LINEAR>;

// TransportVelocityCorrectionCK<SPHBody, Limiter[, BulkParticles]> (transport_velocity_correction_ck.h)
template <int LIMITER = SPHX_LIMITER_NONE, bool BULK_ONLY = false>
class TransportVelocityCorrectionCK
{
  public:
    explicit TransportVelocityCorrectionCK(SPHBody &body, double coefficient = 0.2)
        : body_(body), coefficient_(coefficient)
    {
    }
    void update()
    {
        const Api &a = body_.api();
        a.check(a.sphx_transport_velocity_correction(body_.handle(), coefficient_, LIMITER, BULK_ONLY ? 1 : 0));
    }

  private:
    SPHBody &body_;
    double coefficient_;
};

// FreeSurfaceIndicationCK<Inner<WithUpdate>, Contact<>> (surface_indication_ck.h:170)
class FreeSurfaceIndicationComplexSpatialTemporalCK : public FluidInteraction
{
  public:
    FreeSurfaceIndicationComplexSpatialTemporalCK(Inner &inner) : FluidInteraction(inner, {}) {}
    FreeSurfaceIndicationComplexSpatialTemporalCK(Inner &inner, Contact &wall) : FluidInteraction(inner, {&wall}) {}
    void interact()
    {
        api().check(api().sphx_free_surface_indication(inner_.handle(), walls(), (int)walls_.size()));
    }
};

// reduce dynamics (ReduceDynamicsCK<P, ...>)
class AcousticTimeStepCK
{
  public:
    explicit AcousticTimeStepCK(SPHBody &body, double acoustic_cfl = 0.6) : body_(body), cfl_(acoustic_cfl) {}
    double reduce()
    {
        double dt = 0.0;
        body_.api().check(body_.api().sphx_acoustic_time_step(body_.handle(), cfl_, &dt));
        return dt;
    }

  private:
    SPHBody &body_;
    double cfl_;
};

class AdvectionTimeStepCK
{
  public:
    AdvectionTimeStepCK(SPHBody &body, double u_ref, double advection_cfl = 0.25)
        : body_(body), u_ref_(u_ref), cfl_(advection_cfl)
    {
    }
    double reduce()
    {
        double dt = 0.0;
        body_.api().check(body_.api().sphx_advection_time_step(body_.handle(), u_ref_, cfl_, &dt));
        return dt;
    }

  private:
    SPHBody &body_;
    double u_ref_, cfl_;
};
} // namespace fluid_dynamics

class TotalMechanicalEnergyCK
{
  public:
    TotalMechanicalEnergyCK(SPHBody &body, const std::vector<double> &gravity) : body_(body), g_(gravity) {}
    double reduce()
    {
        double e = 0.0;
        body_.api().check(body_.api().sphx_total_mechanical_energy(body_.handle(), g_.data(), &e));
        return e;
    }

  private:
    SPHBody &body_;
    std::vector<double> g_;
};

class UpperFrontInAxisDirectionCK
{
  public:
    UpperFrontInAxisDirectionCK(SPHBody &body, int axis = 0) : body_(body), axis_(axis) {}
    double reduce()
    {
        double f = 0.0;
        body_.api().check(body_.api().sphx_upper_front_in_axis(body_.handle(), axis_, &f));
        return f;
    }

  private:
    SPHBody &body_;
    int axis_;
};

// Interpolation<Contact<Real>> of a scalar variable of the contact body at the observers
class InterpolationContact
{
  public:
    InterpolationContact(Contact &observer_contact, const std::string &variable)
        : rel_(observer_contact), field_(field_id(variable))
    {
    }
    std::vector<double> interact()
    {
        std::vector<double> out(rel_.getSPHBody().size());
        rel_.getSPHBody().api().check(
            rel_.getSPHBody().api().sphx_interpolate_contact(rel_.handle(), field_, out.data(), out.size()));
        return out;
    }

  private:
    Contact &rel_;
    int field_;
};

// the dynamics wrappers of the reference: StateDynamics, InteractionDynamicsCK,
// ReduceDynamicsCK, each with exec()
template <class ExecutionPolicy, class UpdateType>
class StateDynamics : public UpdateType, public BaseDynamics
{
  public:
    template <class... Args>
    explicit StateDynamics(Args &&...args) : UpdateType(std::forward<Args>(args)...)
    {
    }
    void exec(double = 0.0) { UpdateType::update(); }
};

template <class ExecutionPolicy, class InteractionType>
class InteractionDynamicsCK : public InteractionType, public BaseDynamics
{
  public:
    template <class... Args>
    explicit InteractionDynamicsCK(Args &&...args) : InteractionType(std::forward<Args>(args)...)
    {
    }
    template <class T = InteractionType>
    auto exec(double dt = 0.0) -> decltype(std::declval<T &>().interact(dt))
    {
        return InteractionType::interact(dt);
    }
    template <class T = InteractionType>
    auto exec(double = 0.0) -> decltype(std::declval<T &>().interact())
    {
        return InteractionType::interact();
    }
};

template <class ExecutionPolicy, class ReduceType>
class ReduceDynamicsCK : public ReduceType, public BaseDynamics
{
  public:
    template <class... Args>
    explicit ReduceDynamicsCK(Args &&...args) : ReduceType(std::forward<Args>(args)...)
    {
    }
    double exec(double = 0.0) { return ReduceType::reduce(); }
};

// One advection step of the CK dam-break loop behind one call (see
// sphx_advection_step); returns the acoustic sub-steps taken.
class AdvectionStep
{
  public:
    AdvectionStep(Inner &inner, std::initializer_list<Contact *> walls, CellLinkedList &fluid_cll, double u_ref,
                  int riemann = SPHX_RIEMANN_ACOUSTIC, int flow_type = SPHX_FLOW_FREE_SURFACE,
                  double advection_cfl = 0.25, double acoustic_cfl = 0.6, int correction = SPHX_CORRECTION_NONE,
                  double correction_alpha = 0.5, bool surface_indication = false)
        : inner_(inner), cll_(fluid_cll)
    {
        for (Contact *w : walls)
            walls_.push_back(w->handle());
        cfg_ = {u_ref, advection_cfl, acoustic_cfl, riemann, flow_type, 0, correction, correction_alpha,
                surface_indication ? 1 : 0};
    }
    uint32_t exec(double &physical_time, bool sort_now = false)
    {
        cfg_.sort_now = sort_now ? 1 : 0;
        uint32_t n = 0;
        const Api &a = inner_.getSPHBody().api();
        a.check(a.sphx_advection_step(inner_.handle(), walls_.empty() ? nullptr : walls_.data(), (int)walls_.size(),
                                      cll_.handle(), &cfg_, &physical_time, &n));
        return n;
    }

  private:
    Inner &inner_;
    CellLinkedList &cll_;
    std::vector<sphx_relation *> walls_;
    sphx_advection_config cfg_;
};

} // namespace ck
} // namespace sphx
