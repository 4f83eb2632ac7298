// host_tests.cpp - tests of the C++20 host classes in the style of the
// reference's tests/test_core.cpp (require / require_near /
// require_throws_contains), run on the GPU through the C ABI.
//
//   aurora_host_tests            run every case
//   aurora_host_tests <filter>   run the cases whose name contains filter
//   aurora_host_tests --list     print the case names (no GPU use)
//   aurora_host_tests --cpu <golden dir>   run the host-only cases (no GPU use)
//   aurora_host_tests --aps-signature <file>   print a state file's signature
//   aurora_host_tests --run1d <case.txt> <output dir>   run a 1D case file
//                    through Simulation::run(), writing the reference's outputs
//   aurora_host_tests --run3d <case.txt> <output dir> <32|64>   the same for a
//                    periodic 3D case through Simulation3D::run()
//   aurora_host_tests --run2d <case.txt> <output dir> <32|64>   and for 2D
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "aurora_gpu.hpp"
#include "diagnostics.hpp"
#include "particle_state.hpp"

using namespace pic::gpu;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEV = 1.602176634e-19;
constexpr double kME = 9.1093837139e-31;
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kAmu = 1.66053906660e-27;

void require(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error(what);
}

void require_near(double a, double b, double tol, const std::string& what) {
    if (!(std::abs(a - b) <= tol))
        throw std::runtime_error(what + " (" + std::to_string(a) + " vs " + std::to_string(b) +
                                 ", tolerance " + std::to_string(tol) + ")");
}

template <typename F>
void require_throws_contains(F&& f, const std::string& needle, const std::string& what) {
    try {
        f();
    } catch (const std::exception& e) {
        if (std::string(e.what()).find(needle) != std::string::npos) return;
        throw std::runtime_error(what + ": wrong message: " + e.what());
    }
    throw std::runtime_error(what + ": no exception");
}

using Dev = std::shared_ptr<Device>;

CollisionConfig unit_elastic(double max_frequency = 2.0) {
    CollisionConfig c;
    c.neutral_density = 1.0;
    c.max_frequency = max_frequency;
    c.channels.push_back(CollisionChannelConfig{"elastic", CollisionProcessKind::Elastic,
                                                Table1D{{0.0, 1.0, 2.0}, {0.0, 0.5, 1.0}}});
    return c;
}

// Lotz-shaped synthetic helium-like gas on a log grid (J, m^2)
CollisionConfig synthetic_gas(double density) {
    const int rows = 400;
    std::vector<double> e(rows), el(rows), ex(rows), iz(rows);
    for (int i = 0; i < rows; ++i) {
        const double ev = 1e-3 * std::pow(1e6, static_cast<double>(i) / (rows - 1));
        e[i] = ev * kEV;
        el[i] = 6e-20;
        ex[i] = ev > 19.8 ? 3e-18 * std::log(ev / 19.8) / (ev * 19.8) : 0.0;
        iz[i] = ev > 24.6 ? 6e-18 * std::log(ev / 24.6) / (ev * 24.6) : 0.0;
    }
    CollisionConfig c;
    c.neutral_density = density;
    c.neutral_mass = 4.002602 * kAmu;
    double numax = 0.0;
    for (int i = 0; i < rows; ++i)
        numax = std::max(numax, density * (el[i] + ex[i] + iz[i]) * std::sqrt(2.0 * e[i] / kME));
    c.max_frequency = 1.05 * numax;
    c.channels = {
        {"elastic", CollisionProcessKind::Elastic, Table1D{e, el}},
        {"excitation", CollisionProcessKind::Excitation, Table1D{e, ex}, 19.8 * kEV},
        {"ionization", CollisionProcessKind::Ionization, Table1D{e, iz}, 24.6 * kEV}};
    return c;
}

// ---------------------------------------------------------------- cases
void field_solver_dirichlet_parabola(const Dev& d) {
    // -phi'' = 1 with phi(0) = phi(1) = 0: the central-difference solution
    // is the parabola x (1 - x) / 2 at the nodes (FieldSolver.cpp:733-759)
    Grid g(65, 1.0, Boundary::Dirichlet);
    std::fill(g.rho().begin(), g.rho().end(), 1.0);
    FieldSolver(1.0, d).solve(g, 0.0, 0.0);
    for (size_t i = 0; i < g.nx(); ++i) {
        const double x = g.node_x(i);
        require_near(g.phi()[i], 0.5 * x * (1.0 - x), 1e-12, "Dirichlet parabola");
    }
    require_near(g.electric()[32], 0.0, 1e-12, "Dirichlet field at the centre");
}

void field_solver_3d_periodic_mode(const Dev& d) {
    // test_core.cpp:1360-1438: rho = sin(2 pi x) -> phi = rho / k^2,
    // E_x = -cos(2 pi x) / k with continuous k
    Mesh3D m(32, 16, 8, 1.0, 2.0, 3.0, Boundary::Periodic);
    const double k = 2.0 * kPi;
    for (size_t kk = 0; kk < m.nz(); ++kk)
        for (size_t j = 0; j < m.ny(); ++j)
            for (size_t i = 0; i < m.nx(); ++i)
                m.rho()[m.index(i, j, kk)] = std::sin(k * i * m.dx());
    FieldSolver(2.0, d).solve(m);
    for (size_t i = 0; i < m.nx(); ++i) {
        const size_t n = m.index(i, 3, 5);
        require_near(m.phi()[n], std::sin(k * i * m.dx()) / (2.0 * k * k), 1e-12, "3D phi mode");
        require_near(m.electric_x()[n], -std::cos(k * i * m.dx()) / (2.0 * k), 1e-12,
                     "3D E mode");
        require_near(m.electric_y()[n], 0.0, 1e-12, "3D E_y");
    }
}

void deposit_single_particle_shapes(const Dev& d) {
    // test_core.cpp:1905-1976: trilinear weights of one particle
    Mesh3D m(8, 8, 8, 8.0, 8.0, 8.0, Boundary::Periodic);
    Particle3D p;
    p.position = Vec3{2.25, 3.5, 4.75};
    deposit_charge_cic(m, {p}, 2.0, 3.0, d);
    const double fx[2] = {0.75, 0.25}, fy[2] = {0.5, 0.5}, fz[2] = {0.25, 0.75};
    double total = 0.0;
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
                const double r = m.rho()[m.index(2 + a, 3 + b, 4 + c)];
                require_near(r, 6.0 * fx[a] * fy[b] * fz[c], 1e-12, "trilinear corner weight");
                total += r;
            }
    require_near(total, 6.0, 1e-12, "deposited charge");
    require_near(std::accumulate(m.rho().begin(), m.rho().end(), 0.0), 6.0, 1e-12,
                 "no charge outside the eight corners");
}

void deposit_order_independent(const Dev& d) {
    Mesh2D a(33, 17, 1.0, 1.0, Boundary::Periodic), b = a;
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<Particle2D> ps(20000);
    for (auto& p : ps) p.position = Vec2{u(rng), u(rng)};
    deposit_charge_cic(a, ps, -1.0, 0.25, d);
    std::shuffle(ps.begin(), ps.end(), rng);
    deposit_charge_cic(b, ps, -1.0, 0.25, d);
    require(a.rho() == b.rho(), "fixed-point deposit depends on particle order");
}

void null_collision_elastic_rates(const Dev& d) {
    // test_core.cpp:353-367
    NullCollisionModel model(unit_elastic(), 1.0, 8128, d);
    std::vector<Vec3> v(10000, Vec3{std::sqrt(2.0), 0.0, 0.0});
    const auto s = model.collide(v, 0.1, 0);
    require(s.candidates > 1800 && s.candidates < 2200, "candidate rate");
    require(s.channel_collisions[0] > 620 && s.channel_collisions[0] < 800, "acceptance rate");
    for (const auto& x : v)
        require_near(std::sqrt(x.x * x.x + x.y * x.y + x.z * x.z), std::sqrt(2.0), 1e-5,
                     "elastic speed");
}

void henyey_greenstein_mean_cosine(const Dev& d) {
    // test_core.cpp:397-440
    auto cfg = unit_elastic();
    cfg.channels[0].angular_scattering = AngularScatteringKind::HenyeyGreenstein;
    cfg.channels[0].mean_cosine = Table1D{{0.0, 2.0}, {0.8, 0.8}};
    NullCollisionModel model(cfg, 1.0, 161803, d);
    const float s2 = static_cast<float>(std::sqrt(2.0));
    std::vector<Vec3> v(20000, Vec3{s2, 0.0, 0.0});
    model.collide(v, 0.1, 0);
    double sum = 0.0;
    int n = 0;
    for (const auto& x : v)
        if (x.x != static_cast<double>(s2)) {
            sum += x.x / std::sqrt(2.0);
            ++n;
        }
    require(n > 1200, "anisotropic elastic MCC sampled too few collisions");
    require_near(sum / n, 0.8, 0.04, "Henyey-Greenstein mean cosine");
}

void collision_config_errors(const Dev& d) {
    auto cx = unit_elastic();
    cx.channels[0].process = CollisionProcessKind::ChargeExchange;
    cx.neutral_mass = 40.0;
    require_throws_contains([&] { NullCollisionModel(cx, 39.0, 1, d); },
                            "projectile mass equal to neutral_mass", "mismatched charge exchange");
    auto backward_excitation = unit_elastic(4.0);
    backward_excitation.channels[0].process = CollisionProcessKind::Excitation;
    backward_excitation.channels[0].threshold_energy = 0.5;
    backward_excitation.channels[0].angular_scattering = AngularScatteringKind::Backward;
    require_throws_contains([&] { NullCollisionModel(backward_excitation, 1.0, 1, d); },
                            "supported only for elastic channels", "anisotropic excitation");
    auto dup = unit_elastic();
    dup.channels.push_back(dup.channels[0]);
    require_throws_contains([&] { NullCollisionModel(dup, 1.0, 1, d); },
                            "duplicate MCC collision channel name", "duplicate channel names");
}

void simulation3d_tabulated_field_energy(const Dev& d) {
    // normalised units: n0 = 1, omega_p = 1 for electrons, dx = 1
    Simulation3DConfig cfg;
    cfg.nx = cfg.ny = cfg.nz = 16;
    cfg.length_x = cfg.length_y = cfg.length_z = 16.0;
    cfg.dt = 0.05;
    cfg.steps = 40;
    cfg.permittivity = 1.0;
    const size_t n = 8 * 16 * 16 * 16;
    const double w = 16.0 * 16.0 * 16.0 / static_cast<double>(n);
    cfg.species = {Species3DConfig{"electrons", -1.0, 1.0, w, n, {}, 0.5},
                   Species3DConfig{"ions", 1.0, 400.0, w, n, {}, 0.02}};
    TabulatedVectorField1D b;
    b.axis = 0;
    for (int r = 0; r < 65; ++r) {
        const double x = 16.0 * r / 64.0;
        b.coordinates.push_back(x);
        b.values.push_back(Vec3{0.0, 0.3 * std::exp(-std::pow((x - 8.0) / 2.0, 2)), 0.0});
    }
    cfg.magnetic_field_profile = b;
    Simulation3D sim(cfg, d);
    sim.initialize();
    const auto s0 = sim.sample();
    sim.run();
    const auto s1 = sim.sample();
    require(sim.step_count() == 40, "step count");
    require_near(sim.time(), 40 * 0.05, 1e-12, "simulation time");
    require(s1.live_particles == 2 * n, "periodic run lost particles");
    const double e0 = s0.kinetic_energy + s0.field_energy;
    const double e1 = s1.kinetic_energy + s1.field_energy;
    require_near(e1, e0, 0.02 * e0, "total energy drift over 40 steps");
    const auto ps = sim.particles(0);
    require(ps.size() == n, "particle download");
    for (size_t p = 0; p < ps.size(); p += 997)
        require(ps[p].position.x >= 0.0 && ps[p].position.x <= 16.0, "position inside the box");
}

void simulation1d_discharge_with_ionization(const Dev& d) {
    Config cfg;
    cfg.nx = 65;
    cfg.length = 0.067;
    cfg.boundary = Boundary::Dirichlet;
    cfg.dt = 1.0 / (400.0 * 13.56e6);
    cfg.steps = 100;
    cfg.permittivity = kEps0;
    cfg.phi_right_drive = SinusoidalVoltageConfig{200.0, 13.56e6, 0.0};
    cfg.max_particles_per_species = 20000;
    const double w = 2.56e14 * 0.067 / 3000.0;
    cfg.species = {SpeciesConfig{"electrons", -kEV, kME, w, 3000, 0.0, std::sqrt(8.0 * kEV / kME)},
                   SpeciesConfig{"ions", kEV, 6.67e-27, w, 3000, 0.0,
                                 std::sqrt(0.026 * kEV / 6.67e-27)}};
    cfg.velocity_dimensions = 3;
    cfg.collision_models = {NamedCollisionConfig{"electrons", synthetic_gas(1e22), "electrons",
                                                 "ions"}};
    Simulation sim(cfg, d);
    const RunSummary summary = sim.run();
    require(summary.steps_completed == 100 && summary.final_sample.step == 100, "run summary");
    const auto st = sim.collision_statistics();
    require(st.channel_collisions.size() == 3, "channel statistics");
    require(st.channel_collisions[2] > 0, "no ionization sampled");
    require(st.secondaries == st.channel_collisions[2], "one secondary per ionization");
    const auto e = sim.particles(0);
    const auto i = sim.particles(1);
    require(e.size() == 3000 + st.secondaries && i.size() == 3000 + st.secondaries,
            "products appended to both species");
    const auto loss = sim.boundary_losses(0);
    require(loss.absorbed_left + loss.absorbed_right > 0, "no electron reached a wall");
    require(loss.kinetic_energy_left + loss.kinetic_energy_right > 0.0, "wall energy ledger");
    const Grid g = sim.fields();
    require_near(g.phi()[0], 0.0, 1e-12, "grounded electrode");
}

void swarm_zero_cross_section(const Dev& d) {
    // test_core.cpp:6190-6230
    CollisionConfig gas;
    gas.neutral_density = 1e20;
    gas.max_frequency = 1e7;
    gas.neutral_mass = 4.002602 * kAmu;
    gas.channels = {{"elastic", CollisionProcessKind::Elastic, Table1D{{0.0, 1e-16}, {0.0, 0.0}}}};
    SwarmService svc(gas, d);
    SwarmRequest r{};
    r.reduced_field_td = 1.0;
    r.timestep = 1e-9;
    r.particles = 300;
    r.steps = 40;
    r.burn_in = 10;
    r.blocks = 5;
    r.initial_mean_energy_ev = 0.0;
    r.max_energy_ev = 600.0;
    r.seed = 3;
    const auto res = svc.run_all({r}, 64);
    const double a_dt = kEV * 1e-21 * 1e20 / kME * 1e-9;
    double mean_s = 0.0;
    for (int s = 11; s <= 40; ++s) mean_s += s;
    mean_s /= 30.0;
    require(res[0].status == AURORA_OK, "swarm status");
    require_near(res[0].drift, a_dt * mean_s, 1e-5 * a_dt * mean_s, "zero-sigma drift");
    require(res[0].collisions[0] == 0, "collisions with zero cross section");
}

void swarm_alone_equals_batched(const Dev& d) {
    const auto gas = synthetic_gas(1e21);
    SwarmRequest a{};
    a.reduced_field_td = 150.0;
    a.timestep = 0.099 / gas.max_frequency;
    a.particles = 2000;
    a.steps = 600;
    a.burn_in = 200;
    a.blocks = 8;
    a.initial_mean_energy_ev = 1.0;
    a.max_energy_ev = 1000.0;
    a.seed = 77;
    SwarmRequest b = a;
    b.reduced_field_td = 20.0;
    b.particles = 5000;
    b.seed = 78;
    SwarmService alone(gas, d);
    const auto x = alone.run_all({a}, 600);
    auto d2 = std::make_shared<Device>(0);
    SwarmService mixed(gas, d2);
    const auto y = mixed.run_all({b, a}, 97);
    require(x[0].drift == y[1].drift && x[0].mean_energy_ev == y[1].mean_energy_ev &&
                x[0].ionization_rate == y[1].ionization_rate,
            "a request's result depends on its batch");
}

void error_paths(const Dev& d) {
    require_throws_contains(
        [&] {
            Mesh3D m(8, 8, 8, 1.0, 1.0, 1.0, Boundary::Dirichlet);
            FieldSolver(1.0, d).solve(m);
        },
        "Dirichlet field solves", "3D Dirichlet solve");
    Simulation3DConfig cfg;
    cfg.nx = cfg.ny = cfg.nz = 8;
    cfg.species = {Species3DConfig{}};
    TabulatedVectorField1D b{0, {0.0, 0.5}, {Vec3{}, Vec3{}}};
    cfg.magnetic_field_profile = b;
    require_throws_contains([&] { Simulation3D sim(cfg, d); },
                            "does not cover the simulation domain", "short B table");
    require_throws_contains([&] { Grid g(2, 1.0, Boundary::Periodic); }, "at least 3",
                            "tiny grid");
}

// ---------------------------------------------------------------- .aps
// tests/test_initialization.cpp:786-930 of the reference, restated
std::filesystem::path scratch_dir() {
    auto d = std::filesystem::temp_directory_path() / "aurora_host_tests";
    std::filesystem::create_directories(d);
    return d;
}

void write_state(const std::filesystem::path& p, int dim, int count, const std::string& rows) {
    std::ofstream out(p);
    out << "AuroraPIC-particle-state-v1\ndimension " << dim
        << "\nunits normalized\nweighting species_constant\n"
           "velocity_staggering time_centered\nparticle_count "
        << count << "\nrecords\n"
        << rows << "end\n";
}

void aps_reference_example_signature(const std::filesystem::path& golden) {
    // examples/external_state_1d.cfg pins initial_state_signature for this file
    const auto s = load_external_particle_state(golden / "aps" / "external_state_1d.aps", 4);
    require(s.signature == 15689689587873937355ULL,
            "signature of the reference's example state: " + std::to_string(s.signature));
    require(s.spatial_dimension == 1 && s.particle_count == 4 &&
                s.species.at("electrons").size() == 2 && s.species.at("ions").size() == 2,
            "example state lost metadata or records");
    require(s.species.at("ions")[1].position.x == 0.85 &&
                s.species.at("electrons")[0].velocity.x == 1.0,
            "example state records");
}

void aps_round_trip(const std::filesystem::path&) {
    const auto dir = scratch_dir();
    const auto path = dir / "state.aps";
    write_state(path, 1, 2,
                "particle electrons 0.25 0 0 1.5 0 0\nparticle electrons 0.75 0 0 -0.5 0 0\n");
    const auto s = load_external_particle_state(path, 2);
    require(s.spatial_dimension == 1 && s.particle_count == 2 &&
                s.species.at("electrons").size() == 2 && s.signature != 0,
            "loader lost metadata or records");
    const auto inter = dir / "interleaved.aps";
    write_state(inter, 1, 4,
                "particle electrons 0.1 0 0 1 0 0\nparticle ions 0.2 0 0 2 0 0\n"
                "particle electrons 0.3 0 0 3 0 0\nparticle ions 0.4 0 0 4 0 0\n");
    const auto si = load_external_particle_state(inter, 4);
    require(si.species.at("ions")[1].position.x == 0.4 &&
                si.species.at("electrons")[1].position.x == 0.3,
            "interleaved species records kept in file order per species");
    const auto rt = dir / "roundtrip.aps";
    std::filesystem::remove(rt);
    write_external_particle_state(rt, si);
    require(load_external_particle_state(rt, 4).signature == si.signature,
            "round trip changed the signature");
    require_throws_contains([&] { write_external_particle_state(rt, si); }, "already exists",
                            "writer replaced a file without permission");
    auto mutated = s;
    mutated.species.at("electrons")[0].velocity.x += 0.25;
    require(external_particle_state_signature(mutated) != s.signature,
            "signature ignored a changed record");
    write_external_particle_state(rt, mutated, true);
    require(load_external_particle_state(rt, 2).signature ==
                external_particle_state_signature(mutated),
            "explicit overwrite did not round trip");
    require_throws_contains(
        [&] {
            (void)load_validated_external_particle_state(path, 1, UnitSystem::Normalized,
                                                         {{"electrons", 2}}, "test",
                                                         s.signature + 1);
        },
        "signature mismatch", "signature mismatch accepted");
    validate_external_particle_state(s, 1, UnitSystem::Normalized, {{"electrons", 2}}, "test");
    require_throws_contains([&] { (void)load_external_particle_state(path, 1); },
                            "exceeds the configured limit", "count limit ignored");
    require_throws_contains(
        [&] {
            validate_external_particle_state(s, 2, UnitSystem::Normalized, {{"electrons", 2}},
                                             "test");
        },
        "dimension does not match", "dimension mismatch accepted");
    require_throws_contains(
        [&] {
            validate_external_particle_state(s, 1, UnitSystem::SI, {{"electrons", 2}}, "test");
        },
        "unit system does not match", "unit mismatch accepted");
    require_throws_contains(
        [&] {
            validate_external_particle_state(s, 1, UnitSystem::Normalized, {{"ions", 2}},
                                             "test");
        },
        "missing species 'ions'", "species mismatch accepted");
    const auto meta = dir / "meta.txt";
    write_external_particle_state_metadata(meta, path, {1, 1, UnitSystem::Normalized, 2,
                                                        s.signature}, std::nullopt);
    std::ifstream in(meta);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    require(text.find("signature " + std::to_string(s.signature)) != std::string::npos &&
                text.find("expected_signature none") != std::string::npos,
            "metadata file");
}

// The CSV writers of Simulation.cpp:154-310 on fixed values.
void ledger_csv_writers(const std::filesystem::path&) {
    require(csv_cell("plain") == "plain" && csv_cell("a,b") == "\"a,b\"" &&
                csv_cell("q\"x") == "\"q\"\"x\"",
            "csv_cell quoting");
    std::ostringstream h;
    write_collision_header(h, {"argon.elastic"}, true);
    require(h.str() ==
                "step,time,candidates,null_collisions,collisions_argon.elastic,"
                "tracked_kinetic_energy_change_argon.elastic_J_m-2,cumulative_candidates,"
                "cumulative_null_collisions,cumulative_collisions_argon.elastic,"
                "cumulative_tracked_kinetic_energy_change_argon.elastic_J_m-2\n",
            "collision header");
    CollisionDiagnostics iv{3, 1, {"argon.elastic"}, {2}, {0.1}};
    CollisionDiagnostics tot{7, 2, {"argon.elastic"}, {5}, {1.0 / 3.0}};
    std::ostringstream c;
    write_collision_sample(c, 10, 0.25, iv, tot);
    require(c.str() == "10,0.25,3,1,2,0.10000000000000001,7,2,5,0.33333333333333331\n",
            "collision row: " + c.str());
    std::ostringstream b;
    write_boundary_loss_header(b, {"e,1"}, false);
    require(b.str().find(",\"absorbed_left_count_e,1\",") != std::string::npos,
            "quoted species column");
    std::ostringstream r;
    write_boundary_loss_sample(r, 0, 0.0, 0, {-1.5}, {BoundaryLoss1D{}});
    require(r.str() == "0,0,0,0,0,-0,-0,0,0\n", "boundary-loss row: " + r.str());
    bool threw = false;
    try {
        write_boundary_loss_sample(r, 0, 0.0, 0, {}, {BoundaryLoss1D{}});
    } catch (const std::logic_error&) {
        threw = true;
    }
    require(threw, "species mismatch must throw");
    std::ostringstream pw;
    write_power_transfer_header(pw, {"electrons"}, true);
    write_power_transfer_sample(pw, 5, 1e-9, 0, {SpeciesPower1D{-2.5e-20}});
    require(pw.str() == "step,time,counter_origin_step,electric_work_electrons_J_m-2\n"
                        "5,1.0000000000000001e-09,0,-2.4999999999999999e-20\n",
            "power rows: " + pw.str());
}

void aps_errors(const std::filesystem::path&) {
    const auto dir = scratch_dir();
    const auto p = dir / "bad.aps";
    auto expect = [&](const std::string& text, const std::string& needle) {
        {
            std::ofstream out(p);
            out << text;
        }
        require_throws_contains([&] { (void)load_external_particle_state(p, 10); }, needle,
                                "malformed state file accepted: " + needle);
    };
    const std::string head = "AuroraPIC-particle-state-v1\ndimension 2\nunits normalized\n"
                             "weighting species_constant\nvelocity_staggering time_centered\n";
    expect("AuroraPIC-particle-state-v2\n", "invalid external particle-state magic");
    expect("AuroraPIC-particle-state-v1\ndimension 4\n", "invalid dimension");
    expect("AuroraPIC-particle-state-v1\ndimension 2\nunits cgs\n", "invalid units 'cgs'");
    expect("AuroraPIC-particle-state-v1\ndimension 2\nunits si\nweighting per_particle\n",
           "only weighting species_constant");
    expect(head + "particle_count 1\nrecords\nparticle e 0.1 0.2 0.3 1 2 3\nend\n",
           "nonzero inactive z position");
    expect(head + "particle_count 1\nrecords\nparticle e 0.1 0.2 0 1 2 nan\nend\n",
           "invalid external particle record 0");
    expect(head + "particle_count 1\nrecords\nparticle e 0.1 0.2 0 1 2 3\nend\nextra\n",
           "contains trailing data");
    expect(head + "particle_count 2\nrecords\nparticle e 0.1 0.2 0 1 2 3\nend\n",
           "expected key 'particle'");
    expect(head + "particle_count 0\nrecords\nend\n", "count is zero");
    expect("AuroraPIC-particle-state-v1\ndimension 1\nunits normalized\n"
           "weighting species_constant\nvelocity_staggering time_centered\nparticle_count 1\n"
           "records\nparticle e 0.1 0 0 1 2 0\nend\n",
           "1D1V external particle state has nonzero inactive components");
}

// main.cpp:16-46: the 1e8-update local-run guard and its message
void large_run_guard(const std::filesystem::path&) {
    require(!exceeds_local_run_limit({50'000'000, 50'000'000}, 1), "exactly 1e8 updates");
    require(exceeds_local_run_limit({50'000'000, 50'000'001}, 1), "1e8 + 1 updates");
    require(exceeds_local_run_limit({1'000'000}, 101), "1e6 particles x 101 steps");
    require(!exceeds_local_run_limit({1'000'000}, 0), "zero steps");
    require(exceeds_local_run_limit({std::numeric_limits<std::size_t>::max(), 1}, 1),
            "particle count overflow");
    require_throws_contains([] { require_large_run_acknowledgement(true, false); },
                            "exceed the conservative 100,000,000-update CLI limit",
                            "unacknowledged large run");
    require_large_run_acknowledgement(true, true);
    require_large_run_acknowledgement(false, false);
}

struct CpuCase {
    const char* name;
    void (*body)(const std::filesystem::path&);
};

const CpuCase kCpuCases[] = {
    {"aps_reference_example_signature", aps_reference_example_signature},
    {"aps_round_trip", aps_round_trip},
    {"aps_errors", aps_errors},
    {"ledger_csv_writers", ledger_csv_writers},
    {"large_run_guard", large_run_guard},
};

// The 3D simulation initialised from a state file equals the same records
// set through set_particles; the reference's domain and 1D3V checks.
void simulation3d_external_state(const Dev& d) {
    Simulation3DConfig cfg;
    cfg.nx = cfg.ny = cfg.nz = 16;
    cfg.dt = 0.02;
    cfg.species = {Species3DConfig{"electrons", -1.0, 1.0, 1e-3, 300, {}, 0.0},
                   Species3DConfig{"ions", 1.0, 100.0, 1e-3, 200, {}, 0.0}};
    ExternalParticleState st;
    st.spatial_dimension = 3;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::normal_distribution<double> g(0.0, 0.3);
    for (const auto& s : cfg.species)
        for (size_t i = 0; i < s.particles; ++i)
            st.species[s.name].push_back({{u(rng), u(rng), u(rng)}, {g(rng), g(rng), g(rng)}});
    st.particle_count = 500;
    st.signature = external_particle_state_signature(st);
    const auto path = scratch_dir() / "state3d.aps";
    write_external_particle_state(path, st, true);
    auto a_cfg = cfg;
    a_cfg.initial_state_path = path;
    a_cfg.initial_state_signature = st.signature;
    Simulation3D a(a_cfg, d);
    a.run();
    Simulation3D b(cfg, std::make_shared<Device>(0));
    for (size_t k = 0; k < cfg.species.size(); ++k) {
        std::vector<Particle3D> ps;
        for (const auto& r : st.species.at(cfg.species[k].name))
            ps.push_back({r.position, r.velocity, true, {}});
        b.set_particles(k, ps);
    }
    b.run();
    for (size_t k = 0; k < cfg.species.size(); ++k) {
        const auto pa = a.particles(k), pb = b.particles(k);
        require(pa.size() == pb.size(), "population");
        for (size_t i = 0; i < pa.size(); ++i)
            require(pa[i].position.x == pb[i].position.x && pa[i].velocity_half.z ==
                    pb[i].velocity_half.z, "state-file run differs from set_particles run");
    }
    st.species.at("ions")[0].position.x = 1.0;   // periodic upper end
    st.signature = external_particle_state_signature(st);
    write_external_particle_state(path, st, true);
    require_throws_contains(
        [&] {
            auto c = cfg;
            c.initial_state_path = path;
            Simulation3D s(c, d);
            s.initialize();
        },
        "external particle for species 'ions' lies outside the 3D domain", "domain check");
    require_throws_contains(
        [&] {
            auto c = cfg;
            c.initial_state_path = path;
            c.initial_state_signature = st.signature + 1;
            Simulation3D s(c, d);
            s.initialize();
        },
        "signature mismatch", "pinned signature");
    Config one;
    one.velocity_dimensions = 3;
    one.initial_state_path = path;
    require_throws_contains([&] { Simulation s(one, d); },
                            "not yet available for 1D3V", "1D3V external state");
    one.velocity_dimensions = 2;
    one.initial_state_path.clear();
    require_throws_contains([&] { Simulation s(one, d); },
                            "velocity_dimensions must be 1 or 3", "velocity dimensions");
}

// test_core.cpp:3282-3380: one neutral particle per species leaves through
// each wall in the first step; the ledgers, their CSV contracts and a
// restart from a checkpoint.
void simulation1d_wall_ledgers(const Dev& d) {
    const auto dir = std::filesystem::temp_directory_path() / "aurora_test_1d_boundary_losses";
    std::filesystem::remove_all(dir);
    Config cfg;
    cfg.nx = 3;
    cfg.length = 1.0;
    cfg.dt = 0.2;
    cfg.steps = 1;
    cfg.output_interval = 1;
    cfg.output_dir = dir;
    cfg.boundary = Boundary::Dirichlet;
    cfg.max_particles_per_species = 1;
    cfg.species = {SpeciesConfig{"left", 0.0, 2.0, 3.0, 1, -10.0, 0.0, 64},
                   SpeciesConfig{"right", 0.0, 2.0, 3.0, 1, 10.0, 0.0, 64}};
    Simulation sim(cfg, d);
    const auto summary = sim.run();
    const auto losses = sim.species_boundary_losses();
    require(summary.final_sample.live_particles == 0 && losses.size() == 2 &&
                losses[0].absorbed_left == 1 && losses[0].absorbed_right == 0 &&
                losses[1].absorbed_left == 0 && losses[1].absorbed_right == 1,
            "1D species-resolved wall counts are wrong");
    require_near(losses[0].kinetic_energy_left, 300.0, 1e-12, "left-wall impact energy");
    require_near(losses[1].kinetic_energy_right, 300.0, 1e-12, "right-wall impact energy");
    const auto power = sim.species_power_transfer();
    require_near(power[0].electric_work, 0.0, 1e-14, "neutral particle acquired work");
    require_near(power[1].electric_work, 0.0, 1e-14, "neutral particle acquired work");
    auto text = [](const std::filesystem::path& p) {
        std::ifstream in(p);
        return std::string((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    };
    auto lines = [](const std::string& t) { return std::count(t.begin(), t.end(), '\n'); };
    const std::string bl = text(dir / "boundary_losses.csv");
    require(bl.find("absorbed_left_count_left,absorbed_right_count_left,"
                    "absorbed_left_charge_left_normalized,absorbed_right_charge_left_normalized,"
                    "absorbed_left_kinetic_energy_left_normalized,"
                    "absorbed_right_kinetic_energy_left_normalized") != std::string::npos &&
                lines(bl) == 3,
            "1D wall-loss CSV contract is wrong");
    const std::string pt = text(dir / "power_transfer.csv");
    require(pt.find("step,time,counter_origin_step,electric_work_left_normalized,"
                    "electric_work_right_normalized\n") == 0 &&
                lines(pt) == 3,
            "1D power-transfer CSV contract is wrong");
    const std::string sc = text(dir / "scalars.csv");
    require(sc.find("step,time,kinetic_energy,field_energy,total_energy,charge_l1,"
                    "live_particles,phi_left,phi_right,live_particles_left,"
                    "live_particles_right\n") == 0 && lines(sc) == 3,
            "1D scalar CSV contract is wrong");
    require(std::filesystem::exists(dir / "fields_1.csv"), "final field diagnostics");
    sim.save_checkpoint(dir / "checkpoint_1.ckpt");
    Simulation restarted(cfg, d);
    restarted.load_checkpoint(dir / "checkpoint_1.ckpt");
    const auto rl = restarted.species_boundary_losses();
    require(rl[0].absorbed_left == 1 && rl[1].absorbed_right == 1 &&
                restarted.boundary_loss_origin_step() == 0 && restarted.step_count() == 1,
            "1D checkpoint lost species-resolved wall counts");
    require_near(rl[0].kinetic_energy_left, 300.0, 1e-12, "checkpoint lost wall energy");
    std::filesystem::remove_all(dir);
}

// test_core.cpp:3470-3540: the electric work on collisionless survivors
// equals their kinetic-energy change, and a restart continues it.
void simulation1d_power_transfer(const Dev& d) {
    const auto dir = std::filesystem::temp_directory_path() / "aurora_test_1d_power";
    std::filesystem::remove_all(dir);
    std::filesystem::create_directories(dir);
    Config cfg;
    cfg.nx = 17;
    cfg.length = 1.0;
    cfg.dt = 1e-4;
    cfg.steps = 1;
    cfg.boundary = Boundary::Dirichlet;
    cfg.max_particles_per_species = 32;
    SpeciesConfig e{"electrons", -1.0, 1.0, 1.0, 32, 0.0, 0.0, 64};
    e.init_x_min = 0.1;
    e.init_x_max = 0.9;
    cfg.species = {e};
    Simulation continuous(cfg, d);
    continuous.initialize();
    const double e0 = continuous.sample().kinetic_energy;
    continuous.step();
    const double change = continuous.sample().kinetic_energy - e0;
    const double work = continuous.species_power_transfer()[0].electric_work;
    require(std::abs(work) > 1e-12, "zero-work case");
    require_near(work, change, 1e-12, "survivor work does not equal its energy change");
    continuous.save_checkpoint(dir / "power.ckpt");
    Simulation restarted(cfg, d);
    restarted.load_checkpoint(dir / "power.ckpt");
    require(restarted.power_transfer_origin_step() == 0, "power-transfer origin");
    require_near(restarted.species_power_transfer()[0].electric_work, work, 1e-14,
                 "checkpoint lost cumulative electric work");
    continuous.step();
    restarted.step();
    require_near(restarted.species_power_transfer()[0].electric_work,
                 continuous.species_power_transfer()[0].electric_work, 1e-14,
                 "power transfer is not restart deterministic");
    std::filesystem::remove_all(dir);
}

// run() applies the CLI guard to its configuration: 2 x 600,000 particles
// x 100 steps is 1.2e8 updates
void simulation3d_large_run_guard(const Dev& d) {
    Simulation3DConfig cfg;
    cfg.nx = cfg.ny = cfg.nz = 8;
    cfg.steps = 100;
    cfg.dt = 0.01;
    for (int k = 0; k < 2; ++k) {
        Species3DConfig sc;
        sc.name = k ? "ions" : "electrons";
        sc.charge = k ? 1.0 : -1.0;
        sc.mass = k ? 100.0 : 1.0;
        sc.particles = 600'000;
        sc.weight = 1.0 / 600'000;
        cfg.species.push_back(sc);
    }
    Simulation3D sim(cfg, d);
    require_throws_contains([&] { sim.run(); }, "--allow-large-run I_UNDERSTAND_THIS_IS_A_LARGE_RUN",
                            "large run launched without acknowledgement");
    require(sim.step_count() == 0, "a refused run stepped");
    cfg.allow_large_run = std::string(large_run_acknowledgement);
    Simulation3D lifted(cfg, d);
    lifted.initialize();
    for (int k = 0; k < 98; ++k) lifted.step();   // the run continues from step 98
    const auto r = lifted.run();
    require(r.steps_completed == 100, "acknowledged large run did not complete");
}

struct Case {
    const char* name;
    void (*body)(const Dev&);
};

const Case kCases[] = {
    {"field_solver_dirichlet_parabola", field_solver_dirichlet_parabola},
    {"field_solver_3d_periodic_mode", field_solver_3d_periodic_mode},
    {"deposit_single_particle_shapes", deposit_single_particle_shapes},
    {"deposit_order_independent", deposit_order_independent},
    {"null_collision_elastic_rates", null_collision_elastic_rates},
    {"henyey_greenstein_mean_cosine", henyey_greenstein_mean_cosine},
    {"collision_config_errors", collision_config_errors},
    {"simulation3d_tabulated_field_energy", simulation3d_tabulated_field_energy},
    {"simulation1d_discharge_with_ionization", simulation1d_discharge_with_ionization},
    {"swarm_zero_cross_section", swarm_zero_cross_section},
    {"swarm_alone_equals_batched", swarm_alone_equals_batched},
    {"error_paths", error_paths},
    {"simulation3d_external_state", simulation3d_external_state},
    {"simulation1d_wall_ledgers", simulation1d_wall_ledgers},
    {"simulation1d_power_transfer", simulation1d_power_transfer},
    {"simulation3d_large_run_guard", simulation3d_large_run_guard},
};

}  // namespace

// A 1D case file (written by tests/golden/make_run1d_golden.py): one
// "key values..." line per setting, one "species name charge mass weight
// particles drift thermal x_min x_max" line per species.  FP64 particles.
// A two-column table (energy J, cross section m^2), '#'/';' comments.
Table1D read_table(const std::filesystem::path& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open collision table " + path.string());
    Table1D t;
    std::string line;
    while (std::getline(in, line)) {
        const auto c = line.find_first_of("#;");
        if (c != std::string::npos) line.resize(c);
        std::istringstream ls(line);
        double e = 0.0, s = 0.0;
        if (!(ls >> e)) continue;
        if (!(ls >> s)) throw std::runtime_error("malformed table row in " + path.string());
        t.energies.push_back(e);
        t.values.push_back(s);
    }
    return t;
}

void run1d_case(const std::filesystem::path& case_file, const std::filesystem::path& out) {
    std::ifstream in(case_file);
    if (!in) throw std::runtime_error("cannot open " + case_file.string());
    Config cfg;
    cfg.permittivity = 1.0;
    std::string line;
    size_t most = 1;
    size_t max_particles = 0;
    int precision = 64;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key)) continue;
        if (key == "nx") ls >> cfg.nx;
        else if (key == "precision") ls >> precision;
        else if (key == "max_particles") ls >> max_particles;
        else if (key == "units") {
            std::string u;
            ls >> u;
            cfg.units = u == "si" ? UnitSystem::SI : UnitSystem::Normalized;
            cfg.permittivity = u == "si" ? kEps0 : 1.0;
        } else if (key == "model") {
            // model <name> <target> <density> <max_frequency> <neutral_mass>
            //       <max_candidates> <neutral_temperature K>
            NamedCollisionConfig m;
            ls >> m.name >> m.species >> m.config.neutral_density >> m.config.max_frequency >>
                m.config.neutral_mass >> m.config.max_candidates_per_particle >>
                m.config.neutral_temperature;
            cfg.collision_models.push_back(m);
        } else if (key == "channel") {
            // channel <name> <kind> <threshold J> <table file> <secondary|-> <ion|->
            if (cfg.collision_models.empty()) throw std::runtime_error("channel before model");
            auto& m = cfg.collision_models.back();
            CollisionChannelConfig ch;
            std::string kind, table, sec, ion;
            ls >> ch.name >> kind >> ch.threshold_energy >> table >> sec >> ion;
            ch.process = kind == "elastic"      ? CollisionProcessKind::Elastic
                         : kind == "excitation" ? CollisionProcessKind::Excitation
                         : kind == "ionization" ? CollisionProcessKind::Ionization
                         : kind == "charge_exchange"
                             ? CollisionProcessKind::ChargeExchange
                             : throw std::runtime_error("unknown channel kind " + kind);
            ch.cross_section = read_table(case_file.parent_path() / table);
            if (sec != "-") m.secondary_species = sec;
            if (ion != "-") m.ion_species = ion;
            m.config.channels.push_back(ch);
        }
        else if (key == "length") ls >> cfg.length;
        else if (key == "dt") ls >> cfg.dt;
        else if (key == "steps") ls >> cfg.steps;
        else if (key == "output_interval") ls >> cfg.output_interval;
        else if (key == "velocity_dimensions") ls >> cfg.velocity_dimensions;
        else if (key == "seed") ls >> cfg.seed;
        else if (key == "dirichlet") {
            int v = 0;
            ls >> v;
            cfg.boundary = v ? Boundary::Dirichlet : Boundary::Periodic;
        } else if (key == "electrodes") {
            ls >> cfg.phi_left >> cfg.phi_left_drive.amplitude >> cfg.phi_left_drive.frequency >>
                cfg.phi_left_drive.phase >> cfg.phi_right >> cfg.phi_right_drive.amplitude >>
                cfg.phi_right_drive.frequency >> cfg.phi_right_drive.phase;
        } else if (key == "initial_state") {
            std::string f;
            ls >> f;
            cfg.initial_state_path = case_file.parent_path() / f;
        } else if (key == "species") {
            SpeciesConfig sc;
            ls >> sc.name >> sc.charge >> sc.mass >> sc.weight >> sc.particles >>
                sc.drift_velocity >> sc.thermal_velocity >> sc.init_x_min >> sc.init_x_max;
            most = std::max(most, sc.particles);
            cfg.species.push_back(sc);
        } else {
            throw std::runtime_error("unknown case key " + key);
        }
        if (ls.fail()) throw std::runtime_error("malformed case line: " + line);
    }
    for (auto& sc : cfg.species) sc.precision = precision;
    cfg.max_particles_per_species = std::max(most, max_particles);
    cfg.output_dir = out;
    if (const char* seed = std::getenv("AURORA_CASE_SEED")) cfg.seed = std::stoul(seed);
    Simulation sim(cfg, std::make_shared<Device>(0));
    sim.run();
    // the final live particles, (x, v, v_y, v_z) doubles per species
    for (size_t k = 0; k < cfg.species.size(); ++k) {
        std::ofstream f(out / ("final_" + cfg.species[k].name + ".f64"), std::ios::binary);
        for (const auto& p : sim.particles(k)) {
            if (!p.alive) continue;
            const double row[4] = {p.x, p.v, p.velocity_y, p.velocity_z};
            f.write(reinterpret_cast<const char*>(row), sizeof row);
        }
    }
}

// A periodic 3D case file (tests/golden/make_run3d_golden.py): "n", "length",
// "dt", "steps", "output_interval", "seed", "btable <axis> <file>" (rows x bx
// by bz) and "species name charge mass weight particles thermal" lines.
void run3d_case(const std::filesystem::path& case_file, const std::filesystem::path& out,
                int precision) {
    std::ifstream in(case_file);
    if (!in) throw std::runtime_error("cannot open " + case_file.string());
    Simulation3DConfig cfg;
    cfg.permittivity = 1.0;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key)) continue;
        if (key == "n") ls >> cfg.nx >> cfg.ny >> cfg.nz;
        else if (key == "length") ls >> cfg.length_x >> cfg.length_y >> cfg.length_z;
        else if (key == "dt") ls >> cfg.dt;
        else if (key == "steps") ls >> cfg.steps;
        else if (key == "output_interval") ls >> cfg.output_interval;
        else if (key == "seed") ls >> cfg.seed;
        else if (key == "btable") {
            TabulatedVectorField1D t;
            std::string f;
            ls >> t.axis >> f;
            std::ifstream rows(case_file.parent_path() / f);
            double x, bx, by, bz;
            while (rows >> x >> bx >> by >> bz) {
                t.coordinates.push_back(x);
                t.values.push_back(Vec3{bx, by, bz});
            }
            cfg.magnetic_field_profile = t;
        } else if (key == "species") {
            Species3DConfig sc;
            ls >> sc.name >> sc.charge >> sc.mass >> sc.weight >> sc.particles >>
                sc.thermal_velocity;
            sc.precision = precision;
            cfg.species.push_back(sc);
        } else {
            throw std::runtime_error("unknown case key " + key);
        }
        if (ls.fail()) throw std::runtime_error("malformed case line: " + line);
    }
    cfg.output_dir = out;
    Simulation3D sim(cfg, std::make_shared<Device>(0));
    sim.run();
}

// A periodic 2D case file (tests/golden/make_run2d_golden.py): "n", "length",
// "B" (uniform), "depth", "dt", "steps", "output_interval", "seed" and
// "species" lines as in the 3D case.
void run2d_case(const std::filesystem::path& case_file, const std::filesystem::path& out,
                int precision) {
    std::ifstream in(case_file);
    if (!in) throw std::runtime_error("cannot open " + case_file.string());
    Simulation2DConfig cfg;
    cfg.permittivity = 1.0;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key)) continue;
        if (key == "n") ls >> cfg.nx >> cfg.ny;
        else if (key == "length") ls >> cfg.length_x >> cfg.length_y;
        else if (key == "B") ls >> cfg.magnetic_field.x >> cfg.magnetic_field.y >> cfg.magnetic_field.z;
        else if (key == "depth") ls >> cfg.out_of_plane_depth;
        else if (key == "dt") ls >> cfg.dt;
        else if (key == "steps") ls >> cfg.steps;
        else if (key == "output_interval") ls >> cfg.output_interval;
        else if (key == "seed") ls >> cfg.seed;
        else if (key == "species") {
            Species3DConfig sc;
            ls >> sc.name >> sc.charge >> sc.mass >> sc.weight >> sc.particles >>
                sc.thermal_velocity;
            sc.precision = precision;
            cfg.species.push_back(sc);
        } else {
            throw std::runtime_error("unknown case key " + key);
        }
        if (ls.fail()) throw std::runtime_error("malformed case line: " + line);
    }
    cfg.output_dir = out;
    Simulation2D sim(cfg, std::make_shared<Device>(0));
    sim.run();
}

int main(int argc, char** argv) {
    const std::string arg = argc > 1 ? argv[1] : "";
    if (arg == "--list") {
        for (const auto& c : kCases) std::printf("%s\n", c.name);
        for (const auto& c : kCpuCases) std::printf("%s\n", c.name);
        return 0;
    }
    if (arg == "--aps-signature" && argc > 2) {
        try {
            const auto s = load_external_particle_state(argv[2], std::size_t{1} << 40);
            std::printf("%llu\n", static_cast<unsigned long long>(s.signature));
        } catch (const std::exception& e) {
            std::printf("error: %s\n", e.what());
        }
        return 0;
    }
    if (arg == "--run1d" && argc > 3) {
        try {
            run1d_case(argv[2], argv[3]);
        } catch (const std::exception& e) {
            std::printf("error: %s\n", e.what());
            return 1;
        }
        return 0;
    }
    if (arg == "--run2d" && argc > 4) {
        try {
            run2d_case(argv[2], argv[3], std::atoi(argv[4]));
        } catch (const std::exception& e) {
            std::printf("error: %s\n", e.what());
            return 1;
        }
        return 0;
    }
    if (arg == "--run3d" && argc > 4) {
        try {
            run3d_case(argv[2], argv[3], std::atoi(argv[4]));
        } catch (const std::exception& e) {
            std::printf("error: %s\n", e.what());
            return 1;
        }
        return 0;
    }
    if (arg == "--cpu") {
        const std::filesystem::path golden = argc > 2 ? argv[2] : "tests/golden";
        int passed = 0, failed = 0;
        for (const auto& c : kCpuCases) {
            try {
                c.body(golden);
                std::printf("PASS %s\n", c.name);
                ++passed;
            } catch (const std::exception& e) {
                std::printf("FAIL %s: %s\n", c.name, e.what());
                ++failed;
            }
        }
        std::printf("%d passed, %d failed\n", passed, failed);
        return failed ? 1 : 0;
    }
    int passed = 0, failed = 0;
    for (const auto& c : kCases) {
        if (!arg.empty() && std::string(c.name).find(arg) == std::string::npos) continue;
        std::printf("RUN  %s\n", c.name);
        std::fflush(stdout);
        try {
            auto dev = std::make_shared<Device>(0);
            c.body(dev);
            std::printf("PASS %s\n", c.name);
            ++passed;
        } catch (const std::exception& e) {
            std::printf("FAIL %s: %s\n", c.name, e.what());
            ++failed;
        }
        std::fflush(stdout);
    }
    std::printf("%d passed, %d failed\n", passed, failed);
    return failed ? 1 : 0;
}
