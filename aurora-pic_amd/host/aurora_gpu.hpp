// aurora_gpu.hpp - C++20 host classes of the MI355X particle loop.
//
// They keep the call surface of the reference's hot-path classes
// (include/pic/{Grid,Mesh2D,Mesh3D,FieldSolver,Collision,Simulation,
// Simulation2D,Simulation3D,Swarm}.hpp) for code that selects the GPU
// runtime, and reach the device only through the C ABI of
// include/aurora_hip.h.  Failures raise the reference's exception classes
// with its messages: std::invalid_argument for configuration errors,
// std::overflow_error for capacity, std::out_of_range for tabulated-field
// queries, std::runtime_error for collision-frequency / candidate / energy
// limits and HIP failures.
//
// Differences of substance from the reference's CPU classes:
//   - particle state lives on the device between calls; particles(),
//     fields() and sample() copy back on demand;
//   - NullCollisionModel::collide works on a batch of velocities with one
//     counter-based (Philox4x32-10) stream per particle and step instead of
//     one std::mt19937_64 passed by reference;
//   - species loading (uniform positions, Maxwellian velocities) draws
//     from one std::mt19937_64 in the reference's order (Species.cpp:40-78,
//     Species2D.cpp:29-75, Species3D.cpp:28-76) and so starts from the
//     reference's particles.
#pragma once

#include <array>
#include <cstdint>
#include <filesystem>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "aurora_hip.h"

namespace pic::gpu {

// The reference CLI's local-run guard (main.cpp:16-46): a run whose initial
// particles x steps exceed 1e8 updates is refused unless it is launched
// with this acknowledgement.  run() applies it to its configuration; the
// configs' allow_large_run field carries the acknowledgement, which is how
// a GPU deployment lifts the limit explicitly.
inline constexpr std::string_view large_run_acknowledgement = "I_UNDERSTAND_THIS_IS_A_LARGE_RUN";
inline constexpr std::size_t local_particle_update_limit = 100'000'000;
bool exceeds_local_run_limit(const std::vector<std::size_t>& particles, std::size_t steps);
void require_large_run_acknowledgement(bool exceeds_limit, bool acknowledged);

enum class Boundary { Periodic, Dirichlet };                      // Types.hpp:31
enum class ParticleBoundary { Auto, Absorbing, Reflecting, Periodic };  // Types.hpp:33

struct Vec2 {
    double x{0.0};
    double y{0.0};
};

struct Vec3 {
    double x{0.0};
    double y{0.0};
    double z{0.0};
};

// Units.hpp: which unit contract a run (and an external state file) uses
enum class UnitSystem { Normalized, SI };

// Types.hpp:90-113
struct Particle {
    double x{0.0};
    double v{0.0};
    bool alive{true};
    double v_half{0.0};
    double velocity_y{0.0};
    double velocity_z{0.0};
};

struct Particle2D {
    Vec2 position{};
    Vec2 velocity{};
    bool alive{true};
    Vec2 velocity_half{};
    double velocity_z{0.0};
    double velocity_half_z{0.0};
};

struct Particle3D {
    Vec3 position{};
    Vec3 velocity{};
    bool alive{true};
    Vec3 velocity_half{};
};

// Raises the reference's exception class for an aurora_status.
[[noreturn]] void raise_status(int status, const std::string& message);

// One device context (RuntimePolicy backend = gpu, Runtime.hpp:9).
class Device {
public:
    explicit Device(int ordinal = 0);
    ~Device();
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    aurora_hip_ctx* get() const { return ctx_; }
    void check(int status) const;
    std::string info() const;

private:
    aurora_hip_ctx* ctx_ = nullptr;
};

// ------------------------------------------------------------- meshes
// Grid.hpp: nodes 0..nx-1 over [0, length]; periodic meshes identify the
// ends (the reference's dx = length / nx), Dirichlet meshes have dx =
// length / (nx - 1).
class Grid {
public:
    Grid(std::size_t nx, double length, Boundary boundary);
    std::size_t nx() const { return nx_; }
    double length() const { return length_; }
    double dx() const { return dx_; }
    Boundary boundary() const { return boundary_; }
    std::vector<double>& rho() { return rho_; }
    std::vector<double>& phi() { return phi_; }
    std::vector<double>& electric() { return electric_; }
    const std::vector<double>& rho() const { return rho_; }
    const std::vector<double>& phi() const { return phi_; }
    const std::vector<double>& electric() const { return electric_; }
    void clear_charge();
    double node_x(std::size_t i) const { return static_cast<double>(i) * dx_; }
    double node_volume(std::size_t i) const;

private:
    std::size_t nx_;
    double length_, dx_;
    Boundary boundary_;
    std::vector<double> rho_, phi_, electric_;
};

class Mesh2D {
public:
    Mesh2D(std::size_t nx, std::size_t ny, double length_x, double length_y, Boundary boundary);
    std::size_t nx() const { return nx_; }
    std::size_t ny() const { return ny_; }
    std::size_t size() const { return rho_.size(); }
    double length_x() const { return lx_; }
    double length_y() const { return ly_; }
    double dx() const { return dx_; }
    double dy() const { return dy_; }
    Boundary boundary() const { return boundary_; }
    std::size_t index(std::size_t i, std::size_t j) const { return j * nx_ + i; }
    std::vector<double>& rho() { return rho_; }
    std::vector<double>& phi() { return phi_; }
    std::vector<double>& electric_x() { return ex_; }
    std::vector<double>& electric_y() { return ey_; }
    const std::vector<double>& rho() const { return rho_; }
    const std::vector<double>& phi() const { return phi_; }
    const std::vector<double>& electric_x() const { return ex_; }
    const std::vector<double>& electric_y() const { return ey_; }
    void clear_charge();

private:
    std::size_t nx_, ny_;
    double lx_, ly_, dx_, dy_;
    Boundary boundary_;
    std::vector<double> rho_, phi_, ex_, ey_;
};

class Mesh3D {
public:
    Mesh3D(std::size_t nx, std::size_t ny, std::size_t nz, double length_x, double length_y,
           double length_z, Boundary boundary);
    std::size_t nx() const { return n_[0]; }
    std::size_t ny() const { return n_[1]; }
    std::size_t nz() const { return n_[2]; }
    std::size_t size() const { return rho_.size(); }
    double length_x() const { return l_[0]; }
    double length_y() const { return l_[1]; }
    double length_z() const { return l_[2]; }
    double dx() const { return d_[0]; }
    double dy() const { return d_[1]; }
    double dz() const { return d_[2]; }
    Boundary boundary() const { return boundary_; }
    std::size_t index(std::size_t i, std::size_t j, std::size_t k) const {
        return (k * n_[1] + j) * n_[0] + i;
    }
    std::vector<double>& rho() { return rho_; }
    std::vector<double>& phi() { return phi_; }
    std::vector<double>& electric_x() { return ex_; }
    std::vector<double>& electric_y() { return ey_; }
    std::vector<double>& electric_z() { return ez_; }
    const std::vector<double>& rho() const { return rho_; }
    const std::vector<double>& phi() const { return phi_; }
    const std::vector<double>& electric_x() const { return ex_; }
    const std::vector<double>& electric_y() const { return ey_; }
    const std::vector<double>& electric_z() const { return ez_; }
    void clear_charge();

private:
    std::array<std::size_t, 3> n_;
    std::array<double, 3> l_, d_;
    Boundary boundary_;
    std::vector<double> rho_, phi_, ex_, ey_, ez_;
};

// FieldSolver.hpp:7-22.  Each solve runs on the device: the mesh's rho is
// uploaded, phi and E come back.
class FieldSolver {
public:
    explicit FieldSolver(double permittivity = 1.0, std::shared_ptr<Device> device = nullptr);
    double permittivity() const { return permittivity_; }
    void solve(Grid& grid, double phi_left = 0.0, double phi_right = 0.0) const;
    void solve(Mesh2D& mesh) const;
    void solve(Mesh3D& mesh) const;

private:
    double permittivity_;
    std::shared_ptr<Device> device_;
};

// Species::deposit_charge (Species.cpp:161-185) and deposit_charge_cic
// (Mesh2D.cpp:90-144, Mesh3D.cpp:62-126): adds the particles' charge
// density to the mesh's rho (int64 fixed point on the device, converted per
// node), so the result is independent of particle order.
void deposit_charge_cic(Grid& grid, const std::vector<Particle>& particles, double charge,
                        double weight, const std::shared_ptr<Device>& device = nullptr);
void deposit_charge_cic(Mesh2D& mesh, const std::vector<Particle2D>& particles, double charge,
                        double weight, const std::shared_ptr<Device>& device = nullptr);
void deposit_charge_cic(Mesh3D& mesh, const std::vector<Particle3D>& particles, double charge,
                        double weight, const std::shared_ptr<Device>& device = nullptr);

// ---------------------------------------------------------------- MCC
enum class CollisionProcessKind { Elastic, Excitation, Ionization, Attachment, ChargeExchange };
enum class AngularScatteringKind { Isotropic, Backward, HenyeyGreenstein };
enum class CollisionEnergyFrame { Projectile, CenterOfMass };

// A two-column table as CrossSectionTable / MeanCosineTable hold it after
// reading and scaling (energies in J).
struct Table1D {
    std::vector<double> energies;
    std::vector<double> values;
};

// CollisionChannelConfig (Config.hpp:82-108) with the tables already read.
struct CollisionChannelConfig {
    std::string name{};
    CollisionProcessKind process{CollisionProcessKind::Elastic};
    Table1D cross_section{};
    double threshold_energy{0.0};
    AngularScatteringKind angular_scattering{AngularScatteringKind::Isotropic};
    Table1D mean_cosine{};
    CollisionEnergyFrame energy_frame{CollisionEnergyFrame::Projectile};
};

// CollisionConfig (Config.hpp:110-132), null-collision fields.
struct CollisionConfig {
    double neutral_density{0.0};
    double max_frequency{0.0};
    std::size_t max_candidates_per_particle{64};
    std::vector<CollisionChannelConfig> channels{};
    double neutral_mass{0.0};
    double neutral_temperature{0.0};
};

// CollisionStepStatistics (Collision.hpp:50-63), summed over a batch.
struct CollisionStepStatistics {
    std::uint64_t candidates{0};
    std::uint64_t null_collisions{0};
    std::vector<std::uint64_t> channel_collisions{};
    std::vector<double> channel_projectile_energy_change{};
    std::uint64_t secondaries{0};
};

// NullCollisionModel (Collision.hpp:71-141).  collide() applies the model
// to a batch of velocities in place; particle i of step `step` draws from
// the stream (seed, stream_id, i, step).
class NullCollisionModel {
public:
    NullCollisionModel(CollisionConfig config, double particle_mass, std::uint64_t seed = 1,
                       std::shared_ptr<Device> device = nullptr);
    CollisionStepStatistics collide(std::vector<Vec3>& velocities, double timestep,
                                    std::uint64_t step, std::uint32_t stream_id = 0) const;
    const CollisionConfig& config() const { return config_; }
    const std::vector<std::string>& channel_names() const { return names_; }

private:
    CollisionConfig config_;
    double particle_mass_;
    std::vector<std::string> names_;
    std::shared_ptr<Device> device_;
};

// ---------------------------------------------------------- simulations
// DiagnosticSample / DiagnosticSample2D / DiagnosticSample3D
// (Diagnostics.hpp): FP64 energies from the device (kinetic energy of the
// time-centred velocities, field energy and charge L1 over node volumes),
// the electrode potentials of a 1D Dirichlet run, and the wall counters of
// a 2D/3D run summed over species (faces left, right, bottom, top, back,
// front).
struct DiagnosticSample {
    std::size_t step{0};
    double time{0.0};
    double kinetic_energy{0.0};
    double field_energy{0.0};
    double total_energy{0.0};
    double charge_l1{0.0};
    std::size_t live_particles{0};
    double phi_left{0.0};
    double phi_right{0.0};
    std::vector<std::size_t> live_particles_by_species{};
    std::array<std::uint64_t, 6> absorbed{};
};

// RunSummary (Simulation.hpp): what run() returns.
struct RunSummary {
    std::size_t steps_completed{0};
    double final_time{0.0};
    DiagnosticSample final_sample{};
    bool steady_state_reached{false};
};

// TabulatedVectorField1D (PrescribedField.hpp): rows along one axis.
struct TabulatedVectorField1D {
    int axis{0};                  // 0 x, 1 y, 2 z
    std::vector<double> coordinates{};
    std::vector<Vec3> values{};
};

struct Species3DConfig {          // Species3D.hpp:9-26 (uniform Maxwellian loading)
    std::string name{"electrons"};
    double charge{-1.0};
    double mass{1.0};
    double weight{1.0};
    std::size_t particles{1000};
    Vec3 drift_velocity{};
    double thermal_velocity{0.1};
    std::size_t capacity{0};      // slots (0 = particles)
    int precision{32};            // 64 keeps FP64 offsets and velocities (untiled push)
    double init_x_min{0.0}, init_x_max{-1.0};   // negative maxima mean the full axis
    double init_y_min{0.0}, init_y_max{-1.0};
    double init_z_min{0.0}, init_z_max{-1.0};
};

struct Simulation3DConfig {       // Simulation3D.hpp:26-66, the fields the loop uses
    std::size_t nx{32}, ny{32}, nz{32};
    double length_x{1.0}, length_y{1.0}, length_z{1.0};
    double dt{0.02};
    std::size_t steps{100};
    Boundary boundary{Boundary::Periodic};
    std::array<ParticleBoundary, 6> particle_boundary{ParticleBoundary::Auto, ParticleBoundary::Auto,
                                                      ParticleBoundary::Auto, ParticleBoundary::Auto,
                                                      ParticleBoundary::Auto, ParticleBoundary::Auto};
    Vec3 magnetic_field{};
    std::optional<TabulatedVectorField1D> magnetic_field_profile{};
    unsigned seed{12345};
    double permittivity{1.0};
    std::size_t sort_interval{20};
    std::vector<Species3DConfig> species{};
    // external time-zero state (.aps version 1) instead of Maxwellian loading
    UnitSystem units{UnitSystem::Normalized};
    std::filesystem::path initial_state_path{};
    std::optional<std::uint64_t> initial_state_signature{};
    // run() writes scalars.csv there every output_interval steps (empty: none)
    std::filesystem::path output_dir{};
    std::size_t output_interval{10};
    std::string allow_large_run{};    // large_run_acknowledgement lifts the 1e8-update guard
};

struct Simulation2DConfig {       // Simulation2D.hpp, structured periodic/Dirichlet
    std::size_t nx{64}, ny{64};
    double length_x{1.0}, length_y{1.0};
    double out_of_plane_depth{1.0};
    double dt{0.02};
    std::size_t steps{100};
    Boundary boundary{Boundary::Periodic};
    Vec3 magnetic_field{};
    unsigned seed{12345};
    double permittivity{1.0};
    std::size_t sort_interval{20};
    std::vector<Species3DConfig> species{};
    UnitSystem units{UnitSystem::Normalized};
    std::filesystem::path initial_state_path{};
    std::optional<std::uint64_t> initial_state_signature{};
    std::filesystem::path output_dir{};
    std::size_t output_interval{10};
    std::string allow_large_run{};    // large_run_acknowledgement lifts the 1e8-update guard
};

struct SinusoidalVoltageConfig {  // Config.hpp:139-143
    double amplitude{0.0};
    double frequency{0.0};
    double phase{0.0};
};

// Config::collision_models entry: the model, its target and products.
struct NamedCollisionConfig {
    std::string species{};
    CollisionConfig config{};
    std::string secondary_species{};
    std::string ion_species{};
    // diagnostics name: channel columns are "<name>.<channel>" when set
    std::string name{};
};

struct SpeciesConfig {            // Config.hpp:16-30
    std::string name{"electrons"};
    double charge{-1.0};
    double mass{1.0};
    double weight{1.0};
    std::size_t particles{1000};
    double drift_velocity{0.0};
    double thermal_velocity{0.1};
    int precision{32};            // 64 keeps FP64 offsets and velocities
    double init_x_min{0.0};
    double init_x_max{-1.0};      // negative means the full domain
    double drift_velocity_y{0.0};
    double drift_velocity_z{0.0};
};

struct Config {                   // Config.hpp:170-203, the fields the loop uses
    std::size_t nx{128};
    double length{1.0};
    double dt{0.02};
    std::size_t steps{100};
    Boundary boundary{Boundary::Periodic};
    double phi_left{0.0};
    double phi_right{0.0};
    SinusoidalVoltageConfig phi_left_drive{};
    SinusoidalVoltageConfig phi_right_drive{};
    unsigned seed{12345};
    double permittivity{1.0};
    std::size_t max_particles_per_species{10000000};
    std::vector<NamedCollisionConfig> collision_models{};
    std::vector<SpeciesConfig> species{};
    // 1 (x only) or 3; external state files are 1D1V only, as in the reference
    std::size_t velocity_dimensions{1};
    UnitSystem units{UnitSystem::Normalized};
    std::filesystem::path initial_state_path{};
    std::optional<std::uint64_t> initial_state_signature{};
    // run() writes scalars.csv, fields_<step>.csv, collisions.csv,
    // boundary_losses.csv and power_transfer.csv there (empty: no output)
    std::filesystem::path output_dir{};
    std::size_t output_interval{10};
    std::string allow_large_run{};    // large_run_acknowledgement lifts the 1e8-update guard
};

struct BoundaryLoss1D {           // Simulation.hpp:20-25 (counts and impact energy)
    std::size_t absorbed_left{0};
    std::size_t absorbed_right{0};
    double kinetic_energy_left{0.0};
    double kinetic_energy_right{0.0};
};

struct SpeciesPower1D {           // Simulation.hpp:27-29
    double electric_work{0.0};
};

// CollisionDiagnostics (Collision.hpp:65-71): cumulative since initialize,
// one column per channel ("<model>.<channel>" for named models).
struct CollisionDiagnostics {
    std::uint64_t candidates{0};
    std::uint64_t null_collisions{0};
    std::vector<std::string> channel_names{};
    std::vector<std::uint64_t> channel_collisions{};
    std::vector<double> channel_energy_change{};
};

// Simulation (1D, Simulation.hpp): electrodes, named MCC models with
// ionization products, absorbing walls on Dirichlet meshes.
class Simulation {
public:
    explicit Simulation(Config cfg, std::shared_ptr<Device> device = nullptr);
    void initialize();
    void step();
    // steps until step_count() == steps, writing the outputs of the
    // reference's run() (Simulation.cpp:2902-3057) when output_dir is set
    RunSummary run();
    std::vector<Particle> particles(std::size_t species) const;
    DiagnosticSample sample() const;
    Grid fields() const;
    BoundaryLoss1D boundary_losses(std::size_t species) const;
    std::vector<BoundaryLoss1D> species_boundary_losses() const;
    std::vector<SpeciesPower1D> species_power_transfer() const;
    // the ledgers count from initialize() and survive a checkpoint
    std::size_t boundary_loss_origin_step() const { return 0; }
    std::size_t power_transfer_origin_step() const { return 0; }
    CollisionStepStatistics collision_statistics() const;
    CollisionDiagnostics collision_diagnostics() const;
    // binary checkpoint of the device state behind a physics signature
    void save_checkpoint(const std::filesystem::path& path) const;
    void load_checkpoint(const std::filesystem::path& path);
    double time() const { return time_; }
    std::size_t step_count() const { return step_; }
    const Config& config() const { return cfg_; }

private:
    void sync_clock();
    Config cfg_;
    std::shared_ptr<Device> dev_;
    double time_{0.0};
    std::size_t step_{0};
    bool initialized_{false};
};

class Simulation2D {
public:
    explicit Simulation2D(Simulation2DConfig cfg, std::shared_ptr<Device> device = nullptr);
    void initialize();
    void step();
    RunSummary run();                 // scalars.csv when output_dir is set
    void save_checkpoint(const std::filesystem::path& path) const;
    void load_checkpoint(const std::filesystem::path& path);
    std::vector<Particle2D> particles(std::size_t species) const;
    DiagnosticSample sample() const;
    Mesh2D fields() const;
    double time() const { return time_; }
    std::size_t step_count() const { return step_; }

private:
    Simulation2DConfig cfg_;
    std::shared_ptr<Device> dev_;
    double time_{0.0};
    std::size_t step_{0};
    bool initialized_{false};
};

class Simulation3D {
public:
    explicit Simulation3D(Simulation3DConfig cfg, std::shared_ptr<Device> device = nullptr);
    void initialize();
    void step();
    RunSummary run();                 // scalars.csv when output_dir is set
    void save_checkpoint(const std::filesystem::path& path) const;
    void load_checkpoint(const std::filesystem::path& path);
    // replaces a species' particles (external state), before initialize()
    void set_particles(std::size_t species, const std::vector<Particle3D>& particles);
    std::vector<Particle3D> particles(std::size_t species) const;
    DiagnosticSample sample() const;
    Mesh3D fields() const;
    double time() const { return time_; }
    std::size_t step_count() const { return step_; }

private:
    Simulation3DConfig cfg_;
    std::shared_ptr<Device> dev_;
    double time_{0.0};
    std::size_t step_{0};
    bool initialized_{false};
    std::vector<bool> loaded_;
};

// --------------------------------------------------------------- swarms
using SwarmRequest = aurora_hip_swarm_request;
using SwarmResult = aurora_hip_swarm_result;

// Batched electron swarms (run_field, Swarm.cpp:940-1250): requests are
// admitted against an HBM budget and advanced together.
class SwarmService {
public:
    explicit SwarmService(const CollisionConfig& gas, std::shared_ptr<Device> device = nullptr);
    std::int64_t submit(const SwarmRequest& request, std::uint64_t hbm_budget = 0);
    std::int64_t advance(std::int64_t max_steps);
    std::optional<SwarmResult> poll(std::int64_t handle);
    // every request to completion, in submission order
    std::vector<SwarmResult> run_all(const std::vector<SwarmRequest>& requests,
                                     std::int64_t steps_per_launch = 1000);

private:
    std::shared_ptr<Device> dev_;
};

}  // namespace pic::gpu
