// aurora_gpu.cpp - C++20 host classes over the C ABI (include/aurora_hip.h).
#include "aurora_gpu.hpp"
#include "diagnostics.hpp"
#include "particle_state.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <random>
#include <utility>

namespace pic::gpu {

bool exceeds_local_run_limit(const std::vector<std::size_t>& particles, std::size_t steps) {
    std::size_t total = 0;
    for (std::size_t n : particles) {
        if (n > std::numeric_limits<std::size_t>::max() - total) return true;
        total += n;
    }
    return steps != 0 && total > local_particle_update_limit / steps;
}

void require_large_run_acknowledgement(bool exceeds_limit, bool acknowledged) {
    if (exceeds_limit && !acknowledged)
        throw std::runtime_error(
            "estimated initial particle updates exceed the conservative 100,000,000-update CLI "
            "limit; inspect with --validate-only, or deliberately launch with --allow-large-run "
            "I_UNDERSTAND_THIS_IS_A_LARGE_RUN");
}

namespace {

std::shared_ptr<Device> or_new(std::shared_ptr<Device> d) {
    return d ? std::move(d) : std::make_shared<Device>(0);
}

int code(Boundary b) { return b == Boundary::Periodic ? AURORA_PERIODIC : AURORA_DIRICHLET; }

int policy(ParticleBoundary p, Boundary mesh) {
    switch (p) {
        case ParticleBoundary::Absorbing: return AURORA_PB_ABSORB;
        case ParticleBoundary::Reflecting: return AURORA_PB_REFLECT;
        case ParticleBoundary::Periodic: return AURORA_PB_PERIODIC;
        case ParticleBoundary::Auto: break;
    }
    // Auto: periodic meshes wrap, bounded meshes absorb (Simulation3D.cpp:21-28)
    return mesh == Boundary::Periodic ? AURORA_PB_PERIODIC : AURORA_PB_ABSORB;
}

int kind_code(CollisionProcessKind k) {
    switch (k) {
        case CollisionProcessKind::Elastic: return AURORA_ELASTIC;
        case CollisionProcessKind::Excitation: return AURORA_EXCITATION;
        case CollisionProcessKind::Ionization: return AURORA_IONIZATION;
        case CollisionProcessKind::Attachment: return AURORA_ATTACHMENT;
        case CollisionProcessKind::ChargeExchange: return AURORA_CHARGE_EXCHANGE;
    }
    return -1;
}

// Flattened tables of a CollisionConfig for aurora_hip_mcc_load / swarm_gas.
struct Tables {
    std::vector<int> kind;
    std::vector<double> threshold;
    std::vector<int64_t> rows;
    std::vector<const double*> energy, sigma;
    explicit Tables(const CollisionConfig& c) {
        for (const auto& ch : c.channels) {
            kind.push_back(kind_code(ch.process));
            threshold.push_back(ch.threshold_energy);
            if (ch.cross_section.energies.size() != ch.cross_section.values.size())
                throw std::invalid_argument("collision table columns differ in length");
            rows.push_back(static_cast<int64_t>(ch.cross_section.energies.size()));
            energy.push_back(ch.cross_section.energies.data());
            sigma.push_back(ch.cross_section.values.data());
        }
    }
};

// Channel options of a context model (context = true: the model of that
// index; false: model 0 = the context MCC, 1 = the swarm gas).
void apply_channel_options(const Device& d, int model, const CollisionConfig& c, bool context) {
    for (size_t k = 0; k < c.channels.size(); ++k) {
        const auto& ch = c.channels[k];
        const int ang = static_cast<int>(ch.angular_scattering);
        const int frame = ch.energy_frame == CollisionEnergyFrame::CenterOfMass
                              ? AURORA_FRAME_CENTER_OF_MASS
                              : AURORA_FRAME_PROJECTILE;
        if (ang == 0 && frame == AURORA_FRAME_PROJECTILE && ch.mean_cosine.energies.empty())
            continue;
        const bool hg = ch.angular_scattering == AngularScatteringKind::HenyeyGreenstein;
        if (!hg && !ch.mean_cosine.energies.empty())
            throw std::invalid_argument(
                "mean-cosine data requires angular_model = henyey_greenstein");
        const int64_t rows = hg ? static_cast<int64_t>(ch.mean_cosine.energies.size()) : 0;
        const double* e = hg ? ch.mean_cosine.energies.data() : nullptr;
        const double* v = hg ? ch.mean_cosine.values.data() : nullptr;
        if (context)
            d.check(aurora_hip_mcc_model_options(d.get(), model, static_cast<int>(k), ang, frame,
                                                 rows, e, v));
        else
            d.check(aurora_hip_mcc_channel_options(d.get(), model, static_cast<int>(k), ang,
                                                   frame, rows, e, v));
    }
}

void apply_options(const Device& d, int model, const CollisionConfig& c) {
    apply_channel_options(d, model, c, false);
}

void apply_model_options(const Device& d, int model, const CollisionConfig& c) {
    apply_channel_options(d, model, c, true);
}

void mesh_bind(const Device& d, int dim, const int64_t* n, const double* L, Boundary b,
               double eps, double depth) {
    d.check(aurora_hip_mesh_bind(d.get(), dim, n, L, code(b), eps, depth));
}

// adds the device rho to a host rho vector
void add_rho(const Device& d, std::vector<double>& rho) {
    std::vector<double> r(rho.size());
    d.check(aurora_hip_get_fields(d.get(), r.data(), nullptr, nullptr));
    for (size_t i = 0; i < rho.size(); ++i) rho[i] += r[i];
}

template <typename P>
void upload_and_deposit(const Device& d, int dim, const std::vector<P>& particles, double charge,
                        double weight, std::vector<double>& rho,
                        void (*pos_of)(const P&, double*)) {
    int sid = 0;
    const int64_t n = static_cast<int64_t>(particles.size());
    d.check(aurora_hip_species_add(d.get(), charge, 1.0, weight, std::max<int64_t>(n, 1), 64,
                                   &sid));
    std::vector<double> pos(static_cast<size_t>(n * dim)), vel(static_cast<size_t>(3 * n), 0.0);
    std::vector<uint8_t> alive(static_cast<size_t>(n));
    for (int64_t p = 0; p < n; ++p) {
        pos_of(particles[static_cast<size_t>(p)], pos.data() + dim * p);
        alive[static_cast<size_t>(p)] = particles[static_cast<size_t>(p)].alive ? 1 : 0;
    }
    d.check(aurora_hip_species_upload(d.get(), sid, n, pos.data(), vel.data(), alive.data(),
                                      AURORA_VELOCITY_HALF));
    d.check(aurora_hip_deposit(d.get()));
    add_rho(d, rho);
}

void pos1(const Particle& p, double* x) { x[0] = p.x; }
void pos2(const Particle2D& p, double* x) { x[0] = p.position.x; x[1] = p.position.y; }
// the configured species as the expectations of an external state file
template <typename SpeciesConfigs>
ExternalParticleState load_external(const std::filesystem::path& path, std::size_t dim,
                                    UnitSystem units, const SpeciesConfigs& species,
                                    const std::string& context,
                                    std::optional<std::uint64_t> signature) {
    std::vector<ExternalSpeciesExpectation> expected;
    for (const auto& s : species) expected.push_back({s.name, s.particles});
    return load_validated_external_particle_state(path, dim, units, expected, context,
                                                  signature);
}

void pos3(const Particle3D& p, double* x) {
    x[0] = p.position.x; x[1] = p.position.y; x[2] = p.position.z;
}

struct Downloaded {
    std::vector<double> pos, vh, v;
    std::vector<uint8_t> alive;
    int64_t n = 0;
};

Downloaded download(const Device& d, int sid, int dim, double dt) {
    Downloaded out;
    int64_t slots = 0, live = 0;
    d.check(aurora_hip_species_count(d.get(), sid, &slots, &live));
    const int64_t cap = std::max<int64_t>(slots, 1);
    out.pos.resize(static_cast<size_t>(cap * dim));
    out.vh.resize(static_cast<size_t>(cap * 3));
    out.v.resize(static_cast<size_t>(cap * 3));
    out.alive.resize(static_cast<size_t>(cap));
    d.check(aurora_hip_species_download(d.get(), sid, cap, out.pos.data(), out.vh.data(),
                                        out.alive.data(), nullptr, &out.n));
    d.check(aurora_hip_species_velocity(d.get(), sid, dt, cap, out.v.data()));
    return out;
}

// One diagnostic sample of the context (Diagnostics.cpp:37-59, 114-140,
// 193-229): energies on the device, species kinetic energies added in
// species order, wall counters summed over species, and the electrode
// potentials of a 1D Dirichlet mesh.
DiagnosticSample energies(const Device& d, double dt, size_t nspecies, int dim = 0,
                          Boundary boundary = Boundary::Periodic, std::size_t nx = 0) {
    std::vector<double> ke(std::max<size_t>(nspecies, 1));
    double fe = 0.0, l1 = 0.0;
    int64_t live = 0;
    d.check(aurora_hip_energy(d.get(), dt, ke.data(), &fe, &l1, &live));
    DiagnosticSample s;
    int64_t step = 0;
    d.check(aurora_hip_time(d.get(), &s.time, &step));
    s.step = static_cast<size_t>(step);
    for (size_t k = 0; k < nspecies; ++k) {
        s.kinetic_energy += ke[k];
        int64_t slots = 0, lk = 0;
        d.check(aurora_hip_species_count(d.get(), static_cast<int>(k), &slots, &lk));
        s.live_particles_by_species.push_back(static_cast<size_t>(lk));
        int64_t faces[6] = {};
        d.check(aurora_hip_boundary_losses(d.get(), static_cast<int>(k), faces, nullptr));
        for (int f = 0; f < 6; ++f) s.absorbed[f] += static_cast<uint64_t>(faces[f]);
    }
    s.field_energy = fe;
    s.total_energy = s.kinetic_energy + s.field_energy;
    s.charge_l1 = l1;
    s.live_particles = static_cast<size_t>(live);
    if (dim == 1 && boundary == Boundary::Dirichlet && nx > 0) {
        std::vector<double> phi(nx);
        d.check(aurora_hip_get_fields(d.get(), nullptr, phi.data(), nullptr));
        s.phi_left = phi.front();
        s.phi_right = phi.back();
    }
    return s;
}

// run() of the reference (Simulation*.cpp): samples at step 0, every
// output_interval steps and at the last step.  step_to(k) advances the
// device to step k; write(s) writes one output row set.
template <typename StepTo, typename Write>
RunSummary run_loop(std::size_t start, std::size_t limit, std::size_t interval,
                    StepTo&& step_to, Write&& write) {
    RunSummary summary;
    summary.final_sample = write(true);
    std::size_t step = start;
    while (step < limit) {
        std::size_t next = limit;
        if (interval > 0) next = std::min(limit, (step / interval + 1) * interval);
        step_to(next);
        step = next;
        summary.final_sample = write(false);
    }
    summary.steps_completed = step;
    summary.final_time = summary.final_sample.time;
    return summary;
}

}  // namespace

void raise_status(int status, const std::string& message) {
    switch (status) {
        case AURORA_ERR_INVALID_ARGUMENT:
        case AURORA_ERR_UNSUPPORTED:
            throw std::invalid_argument(message);
        case AURORA_ERR_CAPACITY: throw std::overflow_error(message);
        case AURORA_ERR_B_OUT_OF_RANGE:
        case AURORA_ERR_NOT_FOUND:
            throw std::out_of_range(message);
        default: throw std::runtime_error(message);
    }
}

Device::Device(int ordinal) {
    const int rc = aurora_hip_create(ordinal, &ctx_);
    if (rc != AURORA_OK)
        raise_status(rc == AURORA_ERR_INVALID_ARGUMENT ? AURORA_ERR_HIP : rc,
                     "cannot create a HIP context on device " + std::to_string(ordinal));
}

Device::~Device() {
    if (ctx_) aurora_hip_destroy(ctx_);
}

void Device::check(int status) const {
    if (status != AURORA_OK) raise_status(status, aurora_hip_last_error(ctx_));
}

std::string Device::info() const {
    char buf[256] = {};
    check(aurora_hip_device_info(ctx_, buf, sizeof buf));
    return buf;
}

// -------------------------------------------------------------- meshes
Grid::Grid(std::size_t nx, double length, Boundary boundary)
    : nx_(nx), length_(length), boundary_(boundary) {
    if (nx < 3) throw std::invalid_argument("grid nx must be at least 3");
    if (!std::isfinite(length) || length <= 0.0)
        throw std::invalid_argument("domain length must be positive and finite");
    dx_ = length / static_cast<double>(boundary == Boundary::Periodic ? nx : nx - 1);
    rho_.assign(nx, 0.0);
    phi_.assign(nx, 0.0);
    electric_.assign(nx, 0.0);
}

void Grid::clear_charge() { std::fill(rho_.begin(), rho_.end(), 0.0); }

double Grid::node_volume(std::size_t i) const {
    if (i >= nx_) throw std::out_of_range("grid node index out of range");
    if (boundary_ == Boundary::Periodic) return dx_;
    return (i == 0 || i + 1 == nx_) ? 0.5 * dx_ : dx_;
}

Mesh2D::Mesh2D(std::size_t nx, std::size_t ny, double length_x, double length_y,
               Boundary boundary)
    : nx_(nx), ny_(ny), lx_(length_x), ly_(length_y), boundary_(boundary) {
    if (nx < 3 || ny < 3) throw std::invalid_argument("mesh extents must be at least 3");
    const double cx = static_cast<double>(boundary == Boundary::Periodic ? nx : nx - 1);
    const double cy = static_cast<double>(boundary == Boundary::Periodic ? ny : ny - 1);
    dx_ = lx_ / cx;
    dy_ = ly_ / cy;
    rho_.assign(nx * ny, 0.0);
    phi_.assign(nx * ny, 0.0);
    ex_.assign(nx * ny, 0.0);
    ey_.assign(nx * ny, 0.0);
}

void Mesh2D::clear_charge() { std::fill(rho_.begin(), rho_.end(), 0.0); }

Mesh3D::Mesh3D(std::size_t nx, std::size_t ny, std::size_t nz, double length_x, double length_y,
               double length_z, Boundary boundary)
    : n_{nx, ny, nz}, l_{length_x, length_y, length_z}, boundary_(boundary) {
    for (int a = 0; a < 3; ++a) {
        if (n_[a] < 3) throw std::invalid_argument("mesh extents must be at least 3");
        d_[a] = l_[a] / static_cast<double>(boundary == Boundary::Periodic ? n_[a] : n_[a] - 1);
    }
    const size_t n = nx * ny * nz;
    rho_.assign(n, 0.0);
    phi_.assign(n, 0.0);
    ex_.assign(n, 0.0);
    ey_.assign(n, 0.0);
    ez_.assign(n, 0.0);
}

void Mesh3D::clear_charge() { std::fill(rho_.begin(), rho_.end(), 0.0); }

// --------------------------------------------------------- field solver
FieldSolver::FieldSolver(double permittivity, std::shared_ptr<Device> device)
    : permittivity_(permittivity), device_(std::move(device)) {
    if (!std::isfinite(permittivity) || !(permittivity > 0.0))
        throw std::invalid_argument("relative_permittivity must be positive and finite");
}

void FieldSolver::solve(Grid& grid, double phi_left, double phi_right) const {
    const auto d = or_new(device_);
    const int64_t n = static_cast<int64_t>(grid.nx());
    const double L = grid.length();
    mesh_bind(*d, 1, &n, &L, grid.boundary(), permittivity_, 1.0);
    d->check(aurora_hip_set_rho(d->get(), grid.rho().data()));
    d->check(aurora_hip_field_solve(d->get(), phi_left, phi_right, 1));
    d->check(aurora_hip_get_fields(d->get(), nullptr, grid.phi().data(), grid.electric().data()));
}

void FieldSolver::solve(Mesh2D& mesh) const {
    const auto d = or_new(device_);
    const int64_t n[2] = {static_cast<int64_t>(mesh.nx()), static_cast<int64_t>(mesh.ny())};
    const double L[2] = {mesh.length_x(), mesh.length_y()};
    mesh_bind(*d, 2, n, L, mesh.boundary(), permittivity_, 1.0);
    d->check(aurora_hip_set_rho(d->get(), mesh.rho().data()));
    d->check(aurora_hip_field_solve(d->get(), 0.0, 0.0, 1));
    std::vector<double> E(2 * mesh.size());
    d->check(aurora_hip_get_fields(d->get(), nullptr, mesh.phi().data(), E.data()));
    std::copy(E.begin(), E.begin() + mesh.size(), mesh.electric_x().begin());
    std::copy(E.begin() + mesh.size(), E.end(), mesh.electric_y().begin());
}

void FieldSolver::solve(Mesh3D& mesh) const {
    const auto d = or_new(device_);
    const int64_t n[3] = {static_cast<int64_t>(mesh.nx()), static_cast<int64_t>(mesh.ny()),
                          static_cast<int64_t>(mesh.nz())};
    const double L[3] = {mesh.length_x(), mesh.length_y(), mesh.length_z()};
    mesh_bind(*d, 3, n, L, mesh.boundary(), permittivity_, 1.0);
    d->check(aurora_hip_set_rho(d->get(), mesh.rho().data()));
    d->check(aurora_hip_field_solve(d->get(), 0.0, 0.0, 1));
    const size_t s = mesh.size();
    std::vector<double> E(3 * s);
    d->check(aurora_hip_get_fields(d->get(), nullptr, mesh.phi().data(), E.data()));
    std::copy(E.begin(), E.begin() + s, mesh.electric_x().begin());
    std::copy(E.begin() + s, E.begin() + 2 * s, mesh.electric_y().begin());
    std::copy(E.begin() + 2 * s, E.end(), mesh.electric_z().begin());
}

void deposit_charge_cic(Grid& grid, const std::vector<Particle>& particles, double charge,
                        double weight, const std::shared_ptr<Device>& device) {
    const auto d = or_new(device);
    const int64_t n = static_cast<int64_t>(grid.nx());
    const double L = grid.length();
    mesh_bind(*d, 1, &n, &L, grid.boundary(), 1.0, 1.0);
    upload_and_deposit<Particle>(*d, 1, particles, charge, weight, grid.rho(), pos1);
}

void deposit_charge_cic(Mesh2D& mesh, const std::vector<Particle2D>& particles, double charge,
                        double weight, const std::shared_ptr<Device>& device) {
    const auto d = or_new(device);
    const int64_t n[2] = {static_cast<int64_t>(mesh.nx()), static_cast<int64_t>(mesh.ny())};
    const double L[2] = {mesh.length_x(), mesh.length_y()};
    mesh_bind(*d, 2, n, L, mesh.boundary(), 1.0, 1.0);
    upload_and_deposit<Particle2D>(*d, 2, particles, charge, weight, mesh.rho(), pos2);
}

void deposit_charge_cic(Mesh3D& mesh, const std::vector<Particle3D>& particles, double charge,
                        double weight, const std::shared_ptr<Device>& device) {
    const auto d = or_new(device);
    const int64_t n[3] = {static_cast<int64_t>(mesh.nx()), static_cast<int64_t>(mesh.ny()),
                          static_cast<int64_t>(mesh.nz())};
    const double L[3] = {mesh.length_x(), mesh.length_y(), mesh.length_z()};
    mesh_bind(*d, 3, n, L, mesh.boundary(), 1.0, 1.0);
    upload_and_deposit<Particle3D>(*d, 3, particles, charge, weight, mesh.rho(), pos3);
}

// --------------------------------------------------------------- MCC
NullCollisionModel::NullCollisionModel(CollisionConfig config, double particle_mass,
                                       std::uint64_t seed, std::shared_ptr<Device> device)
    : config_(std::move(config)), particle_mass_(particle_mass), device_(or_new(std::move(device))) {
    for (const auto& ch : config_.channels) {
        if (std::find(names_.begin(), names_.end(), ch.name) != names_.end())
            throw std::invalid_argument("duplicate MCC collision channel name: " + ch.name);
        names_.push_back(ch.name);
    }
    const Device& d = *device_;
    const int64_t n = 3;
    const double L = 1.0;
    mesh_bind(d, 1, &n, &L, Boundary::Periodic, 1.0, 1.0);
    int sid = 0;
    d.check(aurora_hip_species_add(d.get(), -1.0, particle_mass, 1.0, 1, 32, &sid));
    const Tables t(config_);
    const int none[2] = {-1, -1};
    d.check(aurora_hip_mcc_load(d.get(), sid, static_cast<int>(t.kind.size()), t.kind.data(),
                                t.threshold.data(), t.rows.data(), t.energy.data(),
                                t.sigma.data(), config_.neutral_density, config_.max_frequency,
                                static_cast<int64_t>(config_.max_candidates_per_particle),
                                config_.neutral_mass, config_.neutral_temperature, none, seed));
    apply_options(d, 0, config_);
}

CollisionStepStatistics NullCollisionModel::collide(std::vector<Vec3>& velocities,
                                                    double timestep, std::uint64_t step,
                                                    std::uint32_t stream_id) const {
    const size_t nch = config_.channels.size();
    std::vector<double> v(3 * velocities.size());
    for (size_t p = 0; p < velocities.size(); ++p) {
        v[3 * p] = velocities[p].x;
        v[3 * p + 1] = velocities[p].y;
        v[3 * p + 2] = velocities[p].z;
    }
    std::vector<uint64_t> counts(nch + 3, 0);
    std::vector<double> ech(std::max<size_t>(nch, 1), 0.0);
    device_->check(aurora_hip_mcc_collide(device_->get(), static_cast<int64_t>(velocities.size()),
                                          v.data(), timestep, step, stream_id, counts.data(),
                                          ech.data()));
    for (size_t p = 0; p < velocities.size(); ++p)
        velocities[p] = Vec3{v[3 * p], v[3 * p + 1], v[3 * p + 2]};
    CollisionStepStatistics s;
    s.candidates = counts[0];
    s.null_collisions = counts[1];
    s.channel_collisions.assign(counts.begin() + 2, counts.begin() + 2 + static_cast<long>(nch));
    s.channel_projectile_energy_change.assign(ech.begin(), ech.begin() + static_cast<long>(nch));
    s.secondaries = counts[2 + nch];
    return s;
}

// ----------------------------------------------------------- Simulation
Simulation::Simulation(Config cfg, std::shared_ptr<Device> device)
    : cfg_(std::move(cfg)), dev_(or_new(std::move(device))) {
    // Simulation.cpp:341-380
    if (cfg_.initial_state_signature && cfg_.initial_state_path.empty())
        throw std::invalid_argument("initial_state_signature requires initial_state_path");
    if (cfg_.velocity_dimensions != 1 && cfg_.velocity_dimensions != 3)
        throw std::invalid_argument("simulation velocity_dimensions must be 1 or 3");
    if (cfg_.velocity_dimensions == 3 && !cfg_.initial_state_path.empty())
        throw std::invalid_argument(
            "external particle-state initialization is not yet available for 1D3V");
    if (cfg_.max_particles_per_species == 0)
        throw std::invalid_argument("max_particles_per_species must be positive");
    if (cfg_.output_interval == 0) throw std::invalid_argument("output_interval must be positive");
    for (const auto& m : cfg_.collision_models)
        for (const auto& ch : m.config.channels) {
            if (ch.process == CollisionProcessKind::Ionization && cfg_.velocity_dimensions != 3)
                throw std::invalid_argument("1D ionization requires velocity_dimensions = 3");
            if (ch.process == CollisionProcessKind::Attachment)
                throw std::invalid_argument("1D MCC does not support attachment");
        }
    const Device& d = *dev_;
    const int64_t n = static_cast<int64_t>(cfg_.nx);
    mesh_bind(d, 1, &n, &cfg_.length, cfg_.boundary, cfg_.permittivity, 1.0);
    const double left[4] = {cfg_.phi_left, cfg_.phi_left_drive.amplitude,
                            cfg_.phi_left_drive.frequency, cfg_.phi_left_drive.phase};
    const double right[4] = {cfg_.phi_right, cfg_.phi_right_drive.amplitude,
                             cfg_.phi_right_drive.frequency, cfg_.phi_right_drive.phase};
    d.check(aurora_hip_set_electrodes(d.get(), left, right));
    for (const auto& s : cfg_.species) {
        int sid = 0;
        const int64_t cap = static_cast<int64_t>(
            std::max(cfg_.max_particles_per_species, s.particles));
        d.check(aurora_hip_species_add(d.get(), s.charge, s.mass, s.weight, cap, s.precision,
                                       &sid));
    }
    d.check(aurora_hip_set_velocity_dimensions(d.get(),
                                               static_cast<int>(cfg_.velocity_dimensions)));
    auto index_of = [&](const std::string& name) {
        for (size_t k = 0; k < cfg_.species.size(); ++k)
            if (cfg_.species[k].name == name) return static_cast<int>(k);
        throw std::invalid_argument("unknown MCC species: " + name);
    };
    for (const auto& m : cfg_.collision_models) {
        const int target = index_of(m.species);
        int prod[2] = {-1, -1};
        if (!m.secondary_species.empty()) prod[0] = index_of(m.secondary_species);
        if (!m.ion_species.empty()) prod[1] = index_of(m.ion_species);
        const Tables t(m.config);
        // models in config order (Simulation.cpp:490-518): the first through
        // aurora_hip_mcc_load, the others added after it
        int model = 0;
        if (&m == &cfg_.collision_models.front())
            d.check(aurora_hip_mcc_load(
                d.get(), target, static_cast<int>(t.kind.size()), t.kind.data(),
                t.threshold.data(), t.rows.data(), t.energy.data(), t.sigma.data(),
                m.config.neutral_density, m.config.max_frequency,
                static_cast<int64_t>(m.config.max_candidates_per_particle), m.config.neutral_mass,
                m.config.neutral_temperature, prod, cfg_.seed));
        else
            d.check(aurora_hip_mcc_add(
                d.get(), target, static_cast<int>(t.kind.size()), t.kind.data(),
                t.threshold.data(), t.rows.data(), t.energy.data(), t.sigma.data(),
                m.config.neutral_density, m.config.max_frequency,
                static_cast<int64_t>(m.config.max_candidates_per_particle), m.config.neutral_mass,
                m.config.neutral_temperature, prod, cfg_.seed, &model));
        apply_model_options(d, model, m.config);
    }
}

// Simulation::initialize (Simulation.cpp:698-770).  Generated states follow
// Species::initialize for uniform random loading (Species.cpp:40-78): one
// std::mt19937_64 seeded with the config seed, species in config order,
// per particle x, then v_x, then (1D3V) v_y and v_z, each from its own
// distribution object, so the particles equal the reference's.  External
// states (1D1V) take x and v_x from the file.  The device then deposits,
// solves at t = 0 and sets v_half = v - qE dt / 2m.
void Simulation::initialize() {
    const Device& d = *dev_;
    const size_t ns = cfg_.species.size();
    std::vector<std::vector<double>> pos(ns), vel(ns);
    if (!cfg_.initial_state_path.empty()) {
        const auto state = load_external(cfg_.initial_state_path, 1, cfg_.units, cfg_.species,
                                         "1D simulation", cfg_.initial_state_signature);
        for (size_t k = 0; k < ns; ++k) {
            const auto& sc = cfg_.species[k];
            const double lo = sc.init_x_min;
            const double hi = sc.init_x_max < 0.0 ? cfg_.length : sc.init_x_max;
            for (const auto& r : state.species.at(sc.name)) {
                if (r.position.x < lo || r.position.x > hi ||
                    (cfg_.boundary == Boundary::Periodic && r.position.x == cfg_.length))
                    throw std::runtime_error("external particle for species '" + sc.name +
                                             "' lies outside its 1D initialization interval");
                pos[k].push_back(r.position.x);
                vel[k].insert(vel[k].end(), {r.velocity.x, 0.0, 0.0});
            }
        }
    } else {
        std::mt19937_64 rng(cfg_.seed);
        for (size_t k = 0; k < ns; ++k) {
            const auto& sc = cfg_.species[k];
            const double lo = sc.init_x_min;
            const double hi = sc.init_x_max < 0.0 ? cfg_.length : sc.init_x_max;
            std::uniform_real_distribution<double> ux(lo, hi);
            std::normal_distribution<double> nv(sc.drift_velocity, sc.thermal_velocity);
            std::normal_distribution<double> nvy(sc.drift_velocity_y, sc.thermal_velocity);
            std::normal_distribution<double> nvz(sc.drift_velocity_z, sc.thermal_velocity);
            pos[k].reserve(sc.particles);
            vel[k].reserve(3 * sc.particles);
            for (size_t p = 0; p < sc.particles; ++p) {
                double x = ux(rng);
                if (cfg_.boundary == Boundary::Periodic)
                    x = std::fmod(std::fmod(x, cfg_.length) + cfg_.length, cfg_.length);
                else
                    x = std::clamp(x, 0.0, cfg_.length);
                const double vx = nv(rng);
                double vy = 0.0, vz = 0.0;
                if (cfg_.velocity_dimensions == 3) {
                    vy = nvy(rng);
                    vz = nvz(rng);
                }
                pos[k].push_back(x);
                vel[k].insert(vel[k].end(), {vx, vy, vz});
            }
        }
    }
    for (size_t k = 0; k < ns; ++k)
        d.check(aurora_hip_species_upload(d.get(), static_cast<int>(k),
                                          static_cast<int64_t>(pos[k].size()), pos[k].data(),
                                          vel[k].data(), nullptr, AURORA_VELOCITY_CENTRED));
    d.check(aurora_hip_initialize(d.get(), cfg_.dt));
    time_ = 0.0;
    step_ = 0;
    initialized_ = true;
}

void Simulation::sync_clock() {
    int64_t s = 0;
    dev_->check(aurora_hip_time(dev_->get(), &time_, &s));
    step_ = static_cast<size_t>(s);
}

void Simulation::step() {
    if (!initialized_) initialize();
    dev_->check(aurora_hip_step(dev_->get(), cfg_.dt, 1));
    sync_clock();
}

RunSummary Simulation::run() {
    std::vector<std::size_t> counts;
    for (const auto& sc : cfg_.species) counts.push_back(sc.particles);
    require_large_run_acknowledgement(exceeds_local_run_limit(counts, cfg_.steps),
                                      cfg_.allow_large_run == large_run_acknowledgement);
    if (!initialized_) initialize();
    auto step_to = [&](size_t k) {
        dev_->check(aurora_hip_step(dev_->get(), cfg_.dt, static_cast<int64_t>(k - step_)));
        sync_clock();
    };
    std::vector<std::string> names;
    std::vector<double> charge_per_macro;
    for (const auto& sc : cfg_.species) {
        names.push_back(sc.name);
        charge_per_macro.push_back(sc.charge * sc.weight);
    }
    if (cfg_.output_dir.empty())
        return run_loop(step_, cfg_.steps, cfg_.output_interval, step_to,
                        [&](bool) { return sample(); });
    // Simulation.cpp:2942-3050
    const bool si = cfg_.units == UnitSystem::SI;
    Diagnostics diag(cfg_.output_dir, names, cfg_.permittivity);
    diag.write_header();
    const bool collisions = !cfg_.collision_models.empty();
    std::ofstream coll, losses, power;
    CollisionDiagnostics last;
    if (collisions) {
        coll.open(cfg_.output_dir / "collisions.csv");
        if (!coll) throw std::runtime_error("cannot open collision diagnostics output");
        last = collision_diagnostics();
        write_collision_header(coll, last.channel_names, si);
    }
    losses.open(cfg_.output_dir / "boundary_losses.csv");
    if (!losses) throw std::runtime_error("cannot open 1D boundary-loss diagnostics output");
    write_boundary_loss_header(losses, names, si);
    power.open(cfg_.output_dir / "power_transfer.csv");
    if (!power) throw std::runtime_error("cannot open 1D power-transfer diagnostics output");
    write_power_transfer_header(power, names, si);
    auto write = [&](bool first) {
        const DiagnosticSample smp = diag.sample(*this);
        diag.write_sample(smp);
        diag.write_fields(step_, fields());
        if (collisions) {
            // the interval counters start empty at the first row (also after
            // a restart) and hold the change since the previous row after it
            const CollisionDiagnostics now = collision_diagnostics();
            const CollisionDiagnostics& since = first ? now : last;
            CollisionDiagnostics interval = now;
            interval.candidates -= since.candidates;
            interval.null_collisions -= since.null_collisions;
            for (size_t c = 0; c < now.channel_collisions.size(); ++c) {
                interval.channel_collisions[c] -= since.channel_collisions[c];
                interval.channel_energy_change[c] -= since.channel_energy_change[c];
            }
            write_collision_sample(coll, step_, time_, interval, now);
            last = now;
        }
        write_boundary_loss_sample(losses, step_, time_, boundary_loss_origin_step(),
                                   charge_per_macro, species_boundary_losses());
        write_power_transfer_sample(power, step_, time_, power_transfer_origin_step(),
                                    species_power_transfer());
        return smp;
    };
    return run_loop(step_, cfg_.steps, cfg_.output_interval, step_to, write);
}

void Simulation::save_checkpoint(const std::filesystem::path& path) const {
    dev_->check(aurora_hip_checkpoint_save(dev_->get(), path.string().c_str()));
}

void Simulation::load_checkpoint(const std::filesystem::path& path) {
    dev_->check(aurora_hip_checkpoint_load(dev_->get(), path.string().c_str()));
    sync_clock();
    initialized_ = true;
}

std::vector<Particle> Simulation::particles(std::size_t species) const {
    const Downloaded x = download(*dev_, static_cast<int>(species), 1, cfg_.dt);
    std::vector<Particle> out(static_cast<size_t>(x.n));
    for (int64_t p = 0; p < x.n; ++p) {
        Particle& q = out[static_cast<size_t>(p)];
        q.x = x.pos[p];
        q.alive = x.alive[p] != 0;
        q.v_half = x.vh[3 * p];
        q.v = x.v[3 * p];
        q.velocity_y = x.v[3 * p + 1];
        q.velocity_z = x.v[3 * p + 2];
    }
    return out;
}

DiagnosticSample Simulation::sample() const {
    return energies(*dev_, cfg_.dt, cfg_.species.size(), 1, cfg_.boundary, cfg_.nx);
}

Grid Simulation::fields() const {
    Grid g(cfg_.nx, cfg_.length, cfg_.boundary);
    dev_->check(aurora_hip_get_fields(dev_->get(), g.rho().data(), g.phi().data(),
                                      g.electric().data()));
    return g;
}

BoundaryLoss1D Simulation::boundary_losses(std::size_t species) const {
    int64_t counts[6] = {};
    double energy[2] = {};
    dev_->check(aurora_hip_boundary_losses(dev_->get(), static_cast<int>(species), counts, energy));
    return BoundaryLoss1D{static_cast<size_t>(counts[0]), static_cast<size_t>(counts[1]),
                          energy[0], energy[1]};
}

std::vector<BoundaryLoss1D> Simulation::species_boundary_losses() const {
    std::vector<BoundaryLoss1D> out;
    for (size_t k = 0; k < cfg_.species.size(); ++k) out.push_back(boundary_losses(k));
    return out;
}

std::vector<SpeciesPower1D> Simulation::species_power_transfer() const {
    std::vector<SpeciesPower1D> out(cfg_.species.size());
    for (size_t k = 0; k < out.size(); ++k)
        dev_->check(aurora_hip_power_transfer(dev_->get(), static_cast<int>(k),
                                              &out[k].electric_work));
    return out;
}

// One column per channel of every model, in model order (the reference's
// diagnostic offsets, Simulation.cpp:420-480); candidates and nulls summed
// over the models.
CollisionDiagnostics Simulation::collision_diagnostics() const {
    CollisionDiagnostics d;
    for (size_t k = 0; k < cfg_.collision_models.size(); ++k) {
        const NamedCollisionConfig& m = cfg_.collision_models[k];
        const size_t nch = m.config.channels.size();
        for (const auto& ch : m.config.channels)
            d.channel_names.push_back(m.name.empty() ? ch.name : m.name + "." + ch.name);
        std::vector<uint64_t> counts(nch + 3);
        std::vector<double> ech(std::max<size_t>(nch, 1));
        dev_->check(aurora_hip_mcc_model_stats(dev_->get(), static_cast<int>(k), counts.data(),
                                               ech.data()));
        d.candidates += counts[0];
        d.null_collisions += counts[1];
        d.channel_collisions.insert(d.channel_collisions.end(), counts.begin() + 2,
                                    counts.begin() + 2 + static_cast<long>(nch));
        d.channel_energy_change.insert(d.channel_energy_change.end(), ech.begin(),
                                       ech.begin() + static_cast<long>(nch));
    }
    return d;
}

CollisionStepStatistics Simulation::collision_statistics() const {
    CollisionStepStatistics s;
    const CollisionDiagnostics d = collision_diagnostics();
    s.candidates = d.candidates;
    s.null_collisions = d.null_collisions;
    s.channel_collisions = d.channel_collisions;
    s.channel_projectile_energy_change = d.channel_energy_change;
    for (size_t k = 0; k < cfg_.collision_models.size(); ++k) {
        const size_t nch = cfg_.collision_models[k].config.channels.size();
        std::vector<uint64_t> counts(nch + 3);
        dev_->check(aurora_hip_mcc_model_stats(dev_->get(), static_cast<int>(k), counts.data(),
                                               nullptr));
        s.secondaries += counts[2 + nch];
    }
    return s;
}

// --------------------------------------------------------- Simulation2D
Simulation2D::Simulation2D(Simulation2DConfig cfg, std::shared_ptr<Device> device)
    : cfg_(std::move(cfg)), dev_(or_new(std::move(device))) {
    if (cfg_.initial_state_signature && cfg_.initial_state_path.empty())
        throw std::invalid_argument("2D initial_state_signature requires initial_state_path");
    const Device& d = *dev_;
    const int64_t n[2] = {static_cast<int64_t>(cfg_.nx), static_cast<int64_t>(cfg_.ny)};
    const double L[2] = {cfg_.length_x, cfg_.length_y};
    mesh_bind(d, 2, n, L, cfg_.boundary, cfg_.permittivity, cfg_.out_of_plane_depth);
    const double b[3] = {cfg_.magnetic_field.x, cfg_.magnetic_field.y, cfg_.magnetic_field.z};
    d.check(aurora_hip_set_uniform_b(d.get(), b));
    for (const auto& s : cfg_.species) {
        int sid = 0;
        d.check(aurora_hip_species_add(d.get(), s.charge, s.mass, s.weight,
                                       static_cast<int64_t>(std::max(s.capacity, s.particles)),
                                       s.precision, &sid));
    }
    d.check(aurora_hip_set_sort_interval(d.get(), static_cast<int64_t>(cfg_.sort_interval)));
}

void Simulation2D::initialize() {
    const Device& d = *dev_;
    if (!cfg_.initial_state_path.empty()) {
        // Simulation2D.cpp:703-771: every configured species from the file,
        // positions inside the domain (periodic upper ends excluded), then
        // half-step velocities from the self-consistent field
        const auto state = load_external(cfg_.initial_state_path, 2, cfg_.units, cfg_.species,
                                         "2D simulation", cfg_.initial_state_signature);
        for (size_t k = 0; k < cfg_.species.size(); ++k) {
            const auto& records = state.species.at(cfg_.species[k].name);
            const int64_t n = static_cast<int64_t>(records.size());
            std::vector<double> pos(static_cast<size_t>(2 * n)), vel(static_cast<size_t>(3 * n));
            for (int64_t p = 0; p < n; ++p) {
                const auto& r = records[static_cast<size_t>(p)];
                const bool outside =
                    r.position.x < 0.0 || r.position.x > cfg_.length_x || r.position.y < 0.0 ||
                    r.position.y > cfg_.length_y ||
                    (cfg_.boundary == Boundary::Periodic &&
                     (r.position.x == cfg_.length_x || r.position.y == cfg_.length_y));
                if (outside)
                    throw std::runtime_error("external particle for species '" +
                                             cfg_.species[k].name +
                                             "' lies outside the 2D domain");
                pos[2 * p] = r.position.x;
                pos[2 * p + 1] = r.position.y;
                vel[3 * p] = r.velocity.x;
                vel[3 * p + 1] = r.velocity.y;
                vel[3 * p + 2] = r.velocity.z;
            }
            d.check(aurora_hip_species_upload(d.get(), static_cast<int>(k), n, pos.data(),
                                              vel.data(), nullptr, AURORA_VELOCITY_CENTRED));
        }
        if (cfg_.boundary == Boundary::Periodic) d.check(aurora_hip_sort(d.get()));
        d.check(aurora_hip_initialize(d.get(), cfg_.dt));
        time_ = 0.0;
        step_ = 0;
        initialized_ = true;
        return;
    }
    // Species2D::initialize (Species2D.cpp:29-75): one generator for all
    // species in config order; per particle x, y, then v_x, v_y, v_z
    std::mt19937_64 rng(cfg_.seed);
    const double L[2] = {cfg_.length_x, cfg_.length_y};
    for (size_t k = 0; k < cfg_.species.size(); ++k) {
        const auto& s = cfg_.species[k];
        const double hx = s.init_x_max < 0.0 ? L[0] : s.init_x_max;
        const double hy = s.init_y_max < 0.0 ? L[1] : s.init_y_max;
        std::uniform_real_distribution<double> ux(s.init_x_min, hx), uy(s.init_y_min, hy);
        std::normal_distribution<double> vx(s.drift_velocity.x, s.thermal_velocity),
            vy(s.drift_velocity.y, s.thermal_velocity), vz(s.drift_velocity.z, s.thermal_velocity);
        const int64_t n = static_cast<int64_t>(s.particles);
        std::vector<double> pos(static_cast<size_t>(2 * n)), vel(static_cast<size_t>(3 * n));
        for (int64_t p = 0; p < n; ++p) {
            double x[2] = {ux(rng), uy(rng)};
            for (int a = 0; a < 2; ++a)
                x[a] = cfg_.boundary == Boundary::Periodic
                           ? std::fmod(std::fmod(x[a], L[a]) + L[a], L[a])
                           : std::clamp(x[a], 0.0, L[a]);
            pos[2 * p] = x[0];
            pos[2 * p + 1] = x[1];
            vel[3 * p] = vx(rng);
            vel[3 * p + 1] = vy(rng);
            vel[3 * p + 2] = vz(rng);
        }
        d.check(aurora_hip_species_upload(d.get(), static_cast<int>(k), n, pos.data(), vel.data(),
                                          nullptr, AURORA_VELOCITY_CENTRED));
    }
    d.check(aurora_hip_sort(d.get()));
    d.check(aurora_hip_initialize(d.get(), cfg_.dt));
    time_ = 0.0;
    step_ = 0;
    initialized_ = true;
}

void Simulation2D::step() {
    if (!initialized_) initialize();
    dev_->check(aurora_hip_step(dev_->get(), cfg_.dt, 1));
    int64_t s = 0;
    dev_->check(aurora_hip_time(dev_->get(), &time_, &s));
    step_ = static_cast<size_t>(s);
}

RunSummary Simulation2D::run() {
    std::vector<std::size_t> counts;
    for (const auto& sc : cfg_.species) counts.push_back(sc.particles);
    require_large_run_acknowledgement(exceeds_local_run_limit(counts, cfg_.steps),
                                      cfg_.allow_large_run == large_run_acknowledgement);
    if (!initialized_) initialize();
    auto step_to = [&](size_t k) {
        dev_->check(aurora_hip_step(dev_->get(), cfg_.dt, static_cast<int64_t>(k - step_)));
        int64_t s = 0;
        dev_->check(aurora_hip_time(dev_->get(), &time_, &s));
        step_ = static_cast<size_t>(s);
    };
    if (cfg_.output_dir.empty())
        return run_loop(step_, cfg_.steps, cfg_.output_interval, step_to,
                        [&](bool) { return sample(); });
    std::vector<std::string> names;
    for (const auto& sc : cfg_.species) names.push_back(sc.name);
    Diagnostics2D diag(cfg_.output_dir, names, cfg_.permittivity, cfg_.out_of_plane_depth);
    diag.write_header();
    return run_loop(step_, cfg_.steps, cfg_.output_interval, step_to, [&](bool) {
        const DiagnosticSample smp = diag.sample(*this);
        diag.write_sample(smp);
        return smp;
    });
}

void Simulation2D::save_checkpoint(const std::filesystem::path& path) const {
    dev_->check(aurora_hip_checkpoint_save(dev_->get(), path.string().c_str()));
}

void Simulation2D::load_checkpoint(const std::filesystem::path& path) {
    dev_->check(aurora_hip_checkpoint_load(dev_->get(), path.string().c_str()));
    int64_t s = 0;
    dev_->check(aurora_hip_time(dev_->get(), &time_, &s));
    step_ = static_cast<size_t>(s);
    initialized_ = true;
}

std::vector<Particle2D> Simulation2D::particles(std::size_t species) const {
    const Downloaded x = download(*dev_, static_cast<int>(species), 2, cfg_.dt);
    std::vector<Particle2D> out(static_cast<size_t>(x.n));
    for (int64_t p = 0; p < x.n; ++p) {
        Particle2D& q = out[static_cast<size_t>(p)];
        q.position = Vec2{x.pos[2 * p], x.pos[2 * p + 1]};
        q.alive = x.alive[p] != 0;
        q.velocity_half = Vec2{x.vh[3 * p], x.vh[3 * p + 1]};
        q.velocity_half_z = x.vh[3 * p + 2];
        q.velocity = Vec2{x.v[3 * p], x.v[3 * p + 1]};
        q.velocity_z = x.v[3 * p + 2];
    }
    return out;
}

DiagnosticSample Simulation2D::sample() const {
    return energies(*dev_, cfg_.dt, cfg_.species.size(), 2);
}

Mesh2D Simulation2D::fields() const {
    Mesh2D m(cfg_.nx, cfg_.ny, cfg_.length_x, cfg_.length_y, cfg_.boundary);
    std::vector<double> E(2 * m.size());
    dev_->check(aurora_hip_get_fields(dev_->get(), m.rho().data(), m.phi().data(), E.data()));
    std::copy(E.begin(), E.begin() + m.size(), m.electric_x().begin());
    std::copy(E.begin() + m.size(), E.end(), m.electric_y().begin());
    return m;
}

// --------------------------------------------------------- Simulation3D
Simulation3D::Simulation3D(Simulation3DConfig cfg, std::shared_ptr<Device> device)
    : cfg_(std::move(cfg)), dev_(or_new(std::move(device))) {
    if (cfg_.initial_state_signature && cfg_.initial_state_path.empty())
        throw std::invalid_argument("3D initial_state_signature requires initial_state_path");
    const Device& d = *dev_;
    const int64_t n[3] = {static_cast<int64_t>(cfg_.nx), static_cast<int64_t>(cfg_.ny),
                          static_cast<int64_t>(cfg_.nz)};
    const double L[3] = {cfg_.length_x, cfg_.length_y, cfg_.length_z};
    mesh_bind(d, 3, n, L, cfg_.boundary, cfg_.permittivity, 1.0);
    int pol[6];
    for (int f = 0; f < 6; ++f) pol[f] = policy(cfg_.particle_boundary[f], cfg_.boundary);
    d.check(aurora_hip_particle_boundaries(d.get(), pol));
    if (cfg_.magnetic_field_profile) {
        const auto& t = *cfg_.magnetic_field_profile;
        if (t.coordinates.size() != t.values.size())
            throw std::invalid_argument("tabulated field columns differ in length");
        std::vector<double> v;
        for (const Vec3& b : t.values) v.insert(v.end(), {b.x, b.y, b.z});
        d.check(aurora_hip_set_b_profile(d.get(), t.axis, static_cast<int64_t>(t.values.size()),
                                         t.coordinates.data(), v.data()));
    } else {
        const double b[3] = {cfg_.magnetic_field.x, cfg_.magnetic_field.y, cfg_.magnetic_field.z};
        d.check(aurora_hip_set_uniform_b(d.get(), b));
    }
    for (const auto& s : cfg_.species) {
        int sid = 0;
        d.check(aurora_hip_species_add(d.get(), s.charge, s.mass, s.weight,
                                       static_cast<int64_t>(std::max(s.capacity, s.particles)),
                                       s.precision, &sid));
    }
    loaded_.assign(cfg_.species.size(), false);
    d.check(aurora_hip_set_sort_interval(d.get(), static_cast<int64_t>(cfg_.sort_interval)));
}

void Simulation3D::set_particles(std::size_t species, const std::vector<Particle3D>& particles) {
    if (species >= cfg_.species.size()) throw std::out_of_range("unknown species index");
    const int64_t n = static_cast<int64_t>(particles.size());
    std::vector<double> pos(static_cast<size_t>(3 * n)), vel(static_cast<size_t>(3 * n));
    std::vector<uint8_t> alive(static_cast<size_t>(n));
    for (int64_t p = 0; p < n; ++p) {
        const Particle3D& q = particles[static_cast<size_t>(p)];
        pos3(q, pos.data() + 3 * p);
        // time-centred velocities; initialize() derives v_half (Species3D.cpp:72)
        vel[3 * p] = q.velocity.x;
        vel[3 * p + 1] = q.velocity.y;
        vel[3 * p + 2] = q.velocity.z;
        alive[static_cast<size_t>(p)] = q.alive ? 1 : 0;
    }
    dev_->check(aurora_hip_species_upload(dev_->get(), static_cast<int>(species), n, pos.data(),
                                          vel.data(), alive.data(), AURORA_VELOCITY_CENTRED));
    loaded_[species] = true;
}

void Simulation3D::initialize() {
    const Device& d = *dev_;
    if (!cfg_.initial_state_path.empty()) {
        // Simulation3D.cpp:221-290
        const auto state = load_external(cfg_.initial_state_path, 3, cfg_.units, cfg_.species,
                                         "3D simulation", cfg_.initial_state_signature);
        for (size_t k = 0; k < cfg_.species.size(); ++k) {
            const auto& records = state.species.at(cfg_.species[k].name);
            std::vector<Particle3D> ps;
            ps.reserve(records.size());
            for (const auto& r : records) {
                const bool outside =
                    r.position.x < 0.0 || r.position.x > cfg_.length_x || r.position.y < 0.0 ||
                    r.position.y > cfg_.length_y || r.position.z < 0.0 ||
                    r.position.z > cfg_.length_z ||
                    (cfg_.boundary == Boundary::Periodic &&
                     (r.position.x == cfg_.length_x || r.position.y == cfg_.length_y ||
                      r.position.z == cfg_.length_z));
                if (outside)
                    throw std::runtime_error("external particle for species '" +
                                             cfg_.species[k].name +
                                             "' lies outside the 3D domain");
                Particle3D q;
                q.position = r.position;
                q.velocity = r.velocity;
                q.alive = true;
                ps.push_back(q);
            }
            set_particles(k, ps);
        }
    }
    // Species3D::initialize (Species3D.cpp:28-76): one generator for all
    // species in config order; per particle x, y, z, then v_x, v_y, v_z,
    // each from its own distribution object
    std::mt19937_64 rng(cfg_.seed);
    for (size_t k = 0; k < cfg_.species.size(); ++k) {
        if (loaded_[k]) continue;
        const auto& s = cfg_.species[k];
        const double L[3] = {cfg_.length_x, cfg_.length_y, cfg_.length_z};
        const double lo[3] = {s.init_x_min, s.init_y_min, s.init_z_min};
        const double hi[3] = {s.init_x_max < 0.0 ? L[0] : s.init_x_max,
                              s.init_y_max < 0.0 ? L[1] : s.init_y_max,
                              s.init_z_max < 0.0 ? L[2] : s.init_z_max};
        std::uniform_real_distribution<double> ux(lo[0], hi[0]), uy(lo[1], hi[1]),
            uz(lo[2], hi[2]);
        std::normal_distribution<double> vx(s.drift_velocity.x, s.thermal_velocity),
            vy(s.drift_velocity.y, s.thermal_velocity), vz(s.drift_velocity.z, s.thermal_velocity);
        std::vector<Particle3D> ps(s.particles);
        for (auto& q : ps) {
            double x[3] = {ux(rng), uy(rng), uz(rng)};
            for (int a = 0; a < 3; ++a)
                x[a] = cfg_.boundary == Boundary::Periodic
                           ? std::fmod(std::fmod(x[a], L[a]) + L[a], L[a])
                           : std::clamp(x[a], 0.0, L[a]);
            q.position = Vec3{x[0], x[1], x[2]};
            q.velocity.x = vx(rng);
            q.velocity.y = vy(rng);
            q.velocity.z = vz(rng);
            q.alive = true;
        }
        set_particles(k, ps);
    }
    if (cfg_.boundary == Boundary::Periodic) d.check(aurora_hip_sort(d.get()));
    d.check(aurora_hip_initialize(d.get(), cfg_.dt));
    time_ = 0.0;
    step_ = 0;
    initialized_ = true;
}

void Simulation3D::step() {
    if (!initialized_) initialize();
    dev_->check(aurora_hip_step(dev_->get(), cfg_.dt, 1));
    int64_t s = 0;
    dev_->check(aurora_hip_time(dev_->get(), &time_, &s));
    step_ = static_cast<size_t>(s);
}

RunSummary Simulation3D::run() {
    std::vector<std::size_t> counts;
    for (const auto& sc : cfg_.species) counts.push_back(sc.particles);
    require_large_run_acknowledgement(exceeds_local_run_limit(counts, cfg_.steps),
                                      cfg_.allow_large_run == large_run_acknowledgement);
    if (!initialized_) initialize();
    auto step_to = [&](size_t k) {
        dev_->check(aurora_hip_step(dev_->get(), cfg_.dt, static_cast<int64_t>(k - step_)));
        int64_t s = 0;
        dev_->check(aurora_hip_time(dev_->get(), &time_, &s));
        step_ = static_cast<size_t>(s);
    };
    if (cfg_.output_dir.empty())
        return run_loop(step_, cfg_.steps, cfg_.output_interval, step_to,
                        [&](bool) { return sample(); });
    std::vector<std::string> names;
    for (const auto& sc : cfg_.species) names.push_back(sc.name);
    Diagnostics3D diag(cfg_.output_dir, names, cfg_.permittivity);
    diag.write_header();
    return run_loop(step_, cfg_.steps, cfg_.output_interval, step_to, [&](bool) {
        const DiagnosticSample smp = diag.sample(*this);
        diag.write_sample(smp);
        return smp;
    });
}

void Simulation3D::save_checkpoint(const std::filesystem::path& path) const {
    dev_->check(aurora_hip_checkpoint_save(dev_->get(), path.string().c_str()));
}

void Simulation3D::load_checkpoint(const std::filesystem::path& path) {
    dev_->check(aurora_hip_checkpoint_load(dev_->get(), path.string().c_str()));
    int64_t s = 0;
    dev_->check(aurora_hip_time(dev_->get(), &time_, &s));
    step_ = static_cast<size_t>(s);
    initialized_ = true;
}

std::vector<Particle3D> Simulation3D::particles(std::size_t species) const {
    const Downloaded x = download(*dev_, static_cast<int>(species), 3, cfg_.dt);
    std::vector<Particle3D> out(static_cast<size_t>(x.n));
    for (int64_t p = 0; p < x.n; ++p) {
        Particle3D& q = out[static_cast<size_t>(p)];
        q.position = Vec3{x.pos[3 * p], x.pos[3 * p + 1], x.pos[3 * p + 2]};
        q.alive = x.alive[p] != 0;
        q.velocity_half = Vec3{x.vh[3 * p], x.vh[3 * p + 1], x.vh[3 * p + 2]};
        q.velocity = Vec3{x.v[3 * p], x.v[3 * p + 1], x.v[3 * p + 2]};
    }
    return out;
}

DiagnosticSample Simulation3D::sample() const {
    return energies(*dev_, cfg_.dt, cfg_.species.size(), 3);
}

Mesh3D Simulation3D::fields() const {
    Mesh3D m(cfg_.nx, cfg_.ny, cfg_.nz, cfg_.length_x, cfg_.length_y, cfg_.length_z,
             cfg_.boundary);
    const size_t s = m.size();
    std::vector<double> E(3 * s);
    dev_->check(aurora_hip_get_fields(dev_->get(), m.rho().data(), m.phi().data(), E.data()));
    std::copy(E.begin(), E.begin() + s, m.electric_x().begin());
    std::copy(E.begin() + s, E.begin() + 2 * s, m.electric_y().begin());
    std::copy(E.begin() + 2 * s, E.end(), m.electric_z().begin());
    return m;
}

// -------------------------------------------------------------- swarms
SwarmService::SwarmService(const CollisionConfig& gas, std::shared_ptr<Device> device)
    : dev_(or_new(std::move(device))) {
    const Tables t(gas);
    dev_->check(aurora_hip_swarm_gas(dev_->get(), static_cast<int>(t.kind.size()), t.kind.data(),
                                     t.threshold.data(), t.rows.data(), t.energy.data(),
                                     t.sigma.data(), gas.neutral_density, gas.max_frequency,
                                     static_cast<int64_t>(gas.max_candidates_per_particle),
                                     gas.neutral_mass, gas.neutral_temperature));
    apply_options(*dev_, 1, gas);
}

std::int64_t SwarmService::submit(const SwarmRequest& request, std::uint64_t hbm_budget) {
    int64_t h = 0;
    dev_->check(aurora_hip_swarm_submit(dev_->get(), &request, hbm_budget, &h));
    return h;
}

std::int64_t SwarmService::advance(std::int64_t max_steps) {
    int64_t active = 0;
    dev_->check(aurora_hip_swarm_advance(dev_->get(), max_steps, &active));
    return active;
}

std::optional<SwarmResult> SwarmService::poll(std::int64_t handle) {
    int done = 0;
    SwarmResult r{};
    dev_->check(aurora_hip_swarm_poll(dev_->get(), handle, &done, &r));
    if (!done) return std::nullopt;
    return r;
}

std::vector<SwarmResult> SwarmService::run_all(const std::vector<SwarmRequest>& requests,
                                               std::int64_t steps_per_launch) {
    std::vector<int64_t> handles;
    for (const auto& r : requests) handles.push_back(submit(r));
    std::vector<std::optional<SwarmResult>> out(requests.size());
    size_t left = requests.size();
    while (left > 0) {
        advance(steps_per_launch);
        for (size_t k = 0; k < handles.size(); ++k) {
            if (out[k]) continue;
            out[k] = poll(handles[k]);
            if (out[k]) --left;
        }
    }
    std::vector<SwarmResult> res;
    for (auto& o : out) res.push_back(*o);
    return res;
}

}  // namespace pic::gpu
