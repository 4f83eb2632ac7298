// ref_harness.cpp - extern "C" entry points over the REAL reference library
// (AuroraPIC sources compiled unmodified from the reference tree by
// oracle/ref/Makefile into oracle/_ref/).  Used only by tests (to validate the
// oracle restatement and write golden fixtures) and by bench.py's
// cpu_baseline leg (to time the reference's own 3D step on the host).
//
// This file is this repository's code; it contains no reference source.
#include "pic/Collision.hpp"
#include "pic/FieldSolver.hpp"
#include "pic/GasDataset.hpp"
#include "pic/Mesh2D.hpp"
#include "pic/Mesh3D.hpp"
#include "pic/ParticleState.hpp"
#include "pic/PrescribedField.hpp"
#include "pic/Pusher.hpp"
#include "pic/Simulation.hpp"
#include "pic/Simulation2D.hpp"
#include "pic/Simulation3D.hpp"
#include "pic/Species.hpp"
#include "pic/Swarm.hpp"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <random>
#include <sstream>
#include <string>
#include <unistd.h>
#include <vector>

namespace {

pic::Boundary to_bc(int bc) {
    return bc == 0 ? pic::Boundary::Periodic : pic::Boundary::Dirichlet;
}

void set_error(char* err, int len, const std::string& what) {
    if (!err || len <= 0) return;
    std::snprintf(err, static_cast<std::size_t>(len), "%s", what.c_str());
}

std::filesystem::path temp_dir() {
    auto dir = std::filesystem::temp_directory_path() /
               ("aurora_ref_" + std::to_string(::getpid()));
    std::filesystem::create_directories(dir);
    return dir;
}

std::filesystem::path write_table(const std::string& stem, int64_t rows,
                                  const double* e, const double* s) {
    auto path = temp_dir() / (stem + ".dat");
    std::ofstream out(path);
    out << std::setprecision(17);
    for (int64_t i = 0; i < rows; ++i) out << e[i] << ' ' << s[i] << '\n';
    return path;
}

pic::CollisionProcessKind to_kind(int k) {
    switch (k) {
        case 0: return pic::CollisionProcessKind::Elastic;
        case 1: return pic::CollisionProcessKind::Excitation;
        case 2: return pic::CollisionProcessKind::Ionization;
        default: return pic::CollisionProcessKind::Attachment;
    }
}

}  // namespace

extern "C" {

int ref_deposit(int dim, const int64_t* n, const double* L, int bc,
                int64_t np, const double* pos, const uint8_t* alive,
                double charge, double weight, double* rho) {
    try {
        if (dim == 1) {
            pic::Grid grid(n[0], L[0], to_bc(bc));
            pic::SpeciesConfig cfg;
            cfg.charge = charge;
            cfg.weight = weight;
            cfg.particles = static_cast<std::size_t>(np);
            pic::Species sp(cfg, 1);
            sp.particles().resize(static_cast<std::size_t>(np));
            for (int64_t p = 0; p < np; ++p) {
                sp.particles()[p].x = pos[p];
                sp.particles()[p].alive = alive ? alive[p] != 0 : true;
            }
            grid.clear_charge();
            sp.deposit_charge(grid);
            std::memcpy(rho, grid.rho().data(), n[0] * sizeof(double));
        } else if (dim == 2) {
            pic::Mesh2D mesh(n[0], n[1], L[0], L[1], to_bc(bc));
            std::vector<pic::Particle2D> ps(static_cast<std::size_t>(np));
            for (int64_t p = 0; p < np; ++p) {
                ps[p].position = {pos[2 * p], pos[2 * p + 1]};
                ps[p].alive = alive ? alive[p] != 0 : true;
            }
            pic::deposit_charge_cic(mesh, ps, charge, weight);
            std::memcpy(rho, mesh.rho().data(), mesh.size() * sizeof(double));
        } else {
            pic::Mesh3D mesh(n[0], n[1], n[2], L[0], L[1], L[2], to_bc(bc));
            std::vector<pic::Particle3D> ps(static_cast<std::size_t>(np));
            for (int64_t p = 0; p < np; ++p) {
                ps[p].position = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
                ps[p].alive = alive ? alive[p] != 0 : true;
            }
            pic::deposit_charge_cic(mesh, ps, charge, weight);
            std::memcpy(rho, mesh.rho().data(), mesh.size() * sizeof(double));
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_gather(int dim, const int64_t* n, const double* L, int bc,
               int64_t np, const double* pos, const double* E, double* out) {
    try {
        if (dim == 1) {
            pic::Grid grid(n[0], L[0], to_bc(bc));
            std::memcpy(grid.electric().data(), E, n[0] * sizeof(double));
            for (int64_t p = 0; p < np; ++p) out[p] = pic::interpolate_electric(grid, pos[p]);
        } else if (dim == 2) {
            pic::Mesh2D mesh(n[0], n[1], L[0], L[1], to_bc(bc));
            const std::size_t nn = mesh.size();
            std::memcpy(mesh.electric_x().data(), E, nn * sizeof(double));
            std::memcpy(mesh.electric_y().data(), E + nn, nn * sizeof(double));
            for (int64_t p = 0; p < np; ++p) {
                auto e = pic::interpolate_electric(mesh, pic::Vec2{pos[2 * p], pos[2 * p + 1]});
                out[2 * p] = e.x;
                out[2 * p + 1] = e.y;
            }
        } else {
            pic::Mesh3D mesh(n[0], n[1], n[2], L[0], L[1], L[2], to_bc(bc));
            const std::size_t nn = mesh.size();
            std::memcpy(mesh.electric_x().data(), E, nn * sizeof(double));
            std::memcpy(mesh.electric_y().data(), E + nn, nn * sizeof(double));
            std::memcpy(mesh.electric_z().data(), E + 2 * nn, nn * sizeof(double));
            for (int64_t p = 0; p < np; ++p) {
                auto e = pic::interpolate_electric(
                    mesh, pic::Vec3{pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]});
                out[3 * p] = e.x;
                out[3 * p + 1] = e.y;
                out[3 * p + 2] = e.z;
            }
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

// E: dim arrays concatenated; phi: n-node array.
int ref_solve(int dim, const int64_t* n, const double* L, int bc, double eps,
              const double* rho, double phi_left, double phi_right,
              double* phi, double* E) {
    try {
        pic::FieldSolver solver(eps);
        if (dim == 1) {
            pic::Grid grid(n[0], L[0], to_bc(bc));
            std::memcpy(grid.rho().data(), rho, n[0] * sizeof(double));
            solver.solve(grid, phi_left, phi_right);
            std::memcpy(phi, grid.phi().data(), n[0] * sizeof(double));
            std::memcpy(E, grid.electric().data(), n[0] * sizeof(double));
        } else if (dim == 2) {
            pic::Mesh2D mesh(n[0], n[1], L[0], L[1], to_bc(bc));
            const std::size_t nn = mesh.size();
            std::memcpy(mesh.rho().data(), rho, nn * sizeof(double));
            solver.solve(mesh);
            std::memcpy(phi, mesh.phi().data(), nn * sizeof(double));
            std::memcpy(E, mesh.electric_x().data(), nn * sizeof(double));
            std::memcpy(E + nn, mesh.electric_y().data(), nn * sizeof(double));
        } else {
            pic::Mesh3D mesh(n[0], n[1], n[2], L[0], L[1], L[2], to_bc(bc));
            const std::size_t nn = mesh.size();
            std::memcpy(mesh.rho().data(), rho, nn * sizeof(double));
            solver.solve(mesh);
            std::memcpy(phi, mesh.phi().data(), nn * sizeof(double));
            std::memcpy(E, mesh.electric_x().data(), nn * sizeof(double));
            std::memcpy(E + nn, mesh.electric_y().data(), nn * sizeof(double));
            std::memcpy(E + 2 * nn, mesh.electric_z().data(), nn * sizeof(double));
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

// one Boris kick of size h (Pusher.cpp kick_boris) on a 3D particle
void ref_boris(const double* v, const double* E, const double* B, double qm,
               double h, double* out) {
    pic::Particle3D p{};
    p.velocity_half = {v[0], v[1], v[2]};
    pic::kick_boris(p, {E[0], E[1], E[2]}, {B[0], B[1], B[2]}, qm, h);
    out[0] = p.velocity_half.x;
    out[1] = p.velocity_half.y;
    out[2] = p.velocity_half.z;
}

int ref_table_eval(int64_t rows, const double* coords, const double* values,
                   double q, double* out3) {
    try {
        std::vector<double> c(coords, coords + rows);
        std::vector<pic::Vec3> v(static_cast<std::size_t>(rows));
        for (int64_t i = 0; i < rows; ++i)
            v[i] = {values[3 * i], values[3 * i + 1], values[3 * i + 2]};
        pic::TabulatedVectorField1D table(pic::CoordinateAxis::X, c, v);
        auto r = table.evaluate({q, 0.0, 0.0});
        out3[0] = r.x;
        out3[1] = r.y;
        out3[2] = r.z;
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_xs_eval(int64_t rows, const double* e, const double* s, int64_t nq,
                const double* q, double* out) {
    try {
        auto path = write_table("xs", rows, e, s);
        pic::CrossSectionTable table(path);
        for (int64_t i = 0; i < nq; ++i) out[i] = table.evaluate(q[i]);
        return 0;
    } catch (...) {
        return -1;
    }
}

// Runs NullCollisionModel::collide (3V) on np particles with one
// mt19937_64(seed) stream, in particle order.  counts: [candidates, nulls,
// chan0..chanN-1, secondaries]; energy_change: per channel sums.
int ref_mcc_collide(int nchan, const int* kind, const double* threshold,
                    const int64_t* rows, const double* const* energy,
                    const double* const* sigma, double neutral_density,
                    double max_frequency, int64_t max_candidates,
                    double particle_mass, double neutral_mass,
                    double neutral_temperature, int64_t np,
                    const double* v_in, double dt, uint64_t seed,
                    double* v_out, uint64_t* counts, double* energy_change,
                    char* err, int errlen) {
    try {
        pic::CollisionConfig cfg;
        cfg.enabled = true;
        cfg.model = pic::CollisionModelKind::NullCollision;
        cfg.neutral_density = neutral_density;
        cfg.species = "electrons";
        cfg.max_frequency = max_frequency;
        cfg.max_candidates_per_particle = static_cast<std::size_t>(max_candidates);
        cfg.neutral_mass = neutral_mass;
        cfg.neutral_temperature = neutral_temperature;
        if (neutral_temperature > 0.0) cfg.gas_data_units = pic::UnitSystem::SI;
        for (int c = 0; c < nchan; ++c) {
            pic::CollisionChannelConfig ch;
            ch.name = "ch" + std::to_string(c);
            ch.process = to_kind(kind[c]);
            ch.cross_section_file = write_table(ch.name, rows[c], energy[c], sigma[c]);
            ch.threshold_energy = threshold[c];
            if (ch.process == pic::CollisionProcessKind::Ionization) {
                ch.secondary_species = "electrons";
                ch.ion_species = "ions";
            }
            if (ch.process == pic::CollisionProcessKind::Attachment)
                ch.attachment_species = "negative_ions";
            cfg.channels.push_back(ch);
        }
        pic::NullCollisionModel model(cfg, particle_mass);
        std::mt19937_64 rng(seed);
        std::memset(counts, 0, sizeof(uint64_t) * static_cast<std::size_t>(nchan + 3));
        for (int c = 0; c < nchan; ++c) energy_change[c] = 0.0;
        for (int64_t p = 0; p < np; ++p) {
            pic::Vec3 v{v_in[3 * p], v_in[3 * p + 1], v_in[3 * p + 2]};
            auto st = model.collide(v, dt, rng);
            v_out[3 * p] = v.x;
            v_out[3 * p + 1] = v.y;
            v_out[3 * p + 2] = v.z;
            counts[0] += st.candidates;
            counts[1] += st.null_collisions;
            for (int c = 0; c < nchan; ++c) {
                counts[2 + c] += st.channel_collisions[c];
                energy_change[c] += st.channel_projectile_energy_change[c];
            }
            counts[2 + nchan] += st.secondaries.size();
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Swarm scan over one reduced field through run_swarm_benchmark
// (fixed population).  out: drift, drift_se, mean energy, its se, ionization
// rate, its se, candidates, collisions per channel (nchan).
int ref_swarm(int nchan, const int* kind, const double* threshold,
              const int64_t* rows, const double* const* energy,
              const double* const* sigma, double neutral_mass,
              double neutral_density, double max_frequency, double timestep,
              int64_t steps, int64_t burn_in, int64_t particles,
              int64_t blocks, double initial_mean_energy_ev,
              double max_energy_ev, double reduced_field_td, uint64_t seed,
              double* out, char* err, int errlen) {
    try {
        auto dir = temp_dir();
        auto manifest = dir / "swarm.gas";
        {
            std::ofstream m(manifest);
            m << std::setprecision(17);
            m << "gas_data_version = 2\nunits = si\ngas = harness_gas\n"
              << "neutral_mass = " << neutral_mass << "\n"
              << "dataset_id = harness\ndataset_version = 1\n"
              << "data_provenance = harness\ncitation = harness\n"
              << "retrieved = 2026-01-01\nlicense = synthetic\n";
            const char* names[] = {"elastic", "excitation", "ionization", "attachment"};
            for (int c = 0; c < nchan; ++c) {
                auto path = write_table("swarm_ch" + std::to_string(c), rows[c],
                                        energy[c], sigma[c]);
                m << "\n[collision.c" << c << "]\ntype = " << names[kind[c]]
                  << "\ncross_section_file = " << path.string() << "\n";
                if (kind[c] == 1 || kind[c] == 2)
                    m << "threshold_energy = " << threshold[c] << "\n";
            }
        }
        pic::SwarmBenchmarkConfig cfg;
        cfg.gas_data_file = manifest;
        cfg.neutral_density = neutral_density;
        cfg.reduced_fields_td = {reduced_field_td};
        cfg.max_frequency = max_frequency;
        cfg.timestep = timestep;
        cfg.steps = static_cast<std::size_t>(steps);
        cfg.burn_in_steps = static_cast<std::size_t>(burn_in);
        cfg.particles = static_cast<std::size_t>(particles);
        cfg.uncertainty_blocks = static_cast<std::size_t>(blocks);
        cfg.initial_mean_energy_ev = initial_mean_energy_ev;
        cfg.max_energy_ev = max_energy_ev;
        cfg.seed = seed;
        cfg.work_item_limit = 4000000000ULL;
        cfg.output_file = dir / "swarm.csv";
        auto results = pic::run_swarm_benchmark(cfg);
        const auto& r = results.front();
        out[0] = r.electron_drift_velocity_m_s;
        out[1] = r.mean_velocity_x_standard_error_m_s;
        out[2] = r.mean_energy_ev;
        out[3] = r.mean_energy_standard_error_ev;
        out[4] = r.ionization_rate_s;
        out[5] = r.ionization_rate_standard_error_s;
        out[6] = static_cast<double>(r.collision_candidates);
        for (int c = 0; c < nchan; ++c) out[7 + c] = static_cast<double>(r.channels[c].collisions);
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// A .swarm benchmark file through the reference's own loader and driver
// (load_swarm_benchmark_config, Swarm.cpp:1252; run_swarm_benchmark,
// Swarm.cpp:1513).  csv_path (may be empty) receives the reference's own
// result CSV (write_swarm_benchmark_csv).  Per field, out
// receives kSwarmFileCols values: reduced field, drift, drift se, mobility,
// mean energy, its se, ionization rate, its se, max observed energy,
// candidates, nulls, neutral velocity stddev, collisions per channel (up to
// 4, zero-filled), longitudinal and transverse diffusion.  Returns the number
// of fields written.
int ref_swarm_file(const char* swarm_path, int max_fields, double* out, const char* csv_path,
                   char* err, int errlen) {
    constexpr int kCols = 31;
    try {
        auto cfg = pic::load_swarm_benchmark_config(swarm_path);
        cfg.output_file = temp_dir() / "swarm_file.csv";
        auto results = pic::run_swarm_benchmark(cfg);
        if (csv_path && *csv_path)
            pic::write_swarm_benchmark_csv(csv_path, cfg,
                                           pic::load_gas_dataset(cfg.gas_data_file), results);
        const int n = static_cast<int>(std::min<size_t>(results.size(), max_fields));
        for (int f = 0; f < n; ++f) {
            const auto& r = results[f];
            double* o = out + static_cast<size_t>(f) * kCols;
            std::fill(o, o + kCols, 0.0);
            o[0] = r.reduced_field_td;
            o[1] = r.electron_drift_velocity_m_s;
            o[2] = r.mean_velocity_x_standard_error_m_s;
            o[3] = r.reduced_mobility_1_v_m_s;
            o[4] = r.mean_energy_ev;
            o[5] = r.mean_energy_standard_error_ev;
            o[6] = r.ionization_rate_s;
            o[7] = r.ionization_rate_standard_error_s;
            o[8] = r.maximum_observed_energy_ev;
            o[9] = static_cast<double>(r.collision_candidates);
            o[10] = static_cast<double>(r.null_collisions);
            o[11] = r.neutral_velocity_stddev_m_s;
            for (size_t c = 0; c < r.channels.size() && c < 4; ++c)
                o[12 + c] = static_cast<double>(r.channels[c].collisions);
            o[16] = r.longitudinal_diffusion_m2_s;
            o[17] = r.transverse_diffusion_m2_s;
            o[18] = r.final_total_electron_weight;
            o[19] = static_cast<double>(r.final_computational_particles);
            o[20] = r.temporal_growth_rate_s;
            o[21] = r.temporal_growth_rate_standard_error_s;
            o[22] = r.attachment_rate_s;
            o[23] = r.attachment_rate_standard_error_s;
            o[24] = r.net_creation_rate_s;
            o[25] = r.net_creation_rate_standard_error_s;
            o[26] = r.townsend_available ? 1.0 : 0.0;
            o[27] = r.growth_over_flux_drift_townsend_1_m;
            o[28] = r.growth_over_flux_drift_townsend_standard_error_1_m;
            o[29] = r.rate_balance_effective_townsend_1_m;
            o[30] = r.rate_balance_effective_townsend_standard_error_1_m;
        }
        return n;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// The reference's own .aps reader on a file: its canonical signature, or
// its exception text (return -1).
int ref_aps_signature(const char* path, uint64_t* signature, char* err, int errlen) {
    try {
        const auto state = pic::load_external_particle_state(path, std::size_t{1} << 40);
        *signature = state.signature;
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Times the reference's own Simulation3D::step() on a periodic mesh with
// species initialised by the reference's initializer.  species: nspecies
// rows of (charge, mass, weight, particles, thermal_velocity).  B profile
// along x (brows rows) or none.  Returns seconds per step (wall).
int ref_sim3d_time(const int64_t* n, const double* L, double dt,
                   double permittivity_rel, int si_units, int nspecies,
                   const double* species, int64_t brows,
                   const double* bcoords, const double* bvalues,
                   int steps, double* seconds_per_step,
                   int64_t* live_particles, char* err, int errlen) {
    try {
        pic::Simulation3DConfig cfg;
        cfg.units.system = si_units ? pic::UnitSystem::SI : pic::UnitSystem::Normalized;
        cfg.units.relative_permittivity = permittivity_rel;
        cfg.nx = n[0];
        cfg.ny = n[1];
        cfg.nz = n[2];
        cfg.length_x = L[0];
        cfg.length_y = L[1];
        cfg.length_z = L[2];
        cfg.dt = dt;
        cfg.steps = static_cast<std::size_t>(steps);
        cfg.output_interval = 1000000;
        cfg.boundary = pic::Boundary::Periodic;
        if (brows > 0) {
            std::vector<double> c(bcoords, bcoords + brows);
            std::vector<pic::Vec3> v(static_cast<std::size_t>(brows));
            for (int64_t i = 0; i < brows; ++i)
                v[i] = {bvalues[3 * i], bvalues[3 * i + 1], bvalues[3 * i + 2]};
            cfg.magnetic_field_profile = pic::TabulatedVectorField1D(pic::CoordinateAxis::X, c, v);
        }
        for (int s = 0; s < nspecies; ++s) {
            pic::Species3DConfig sc;
            sc.name = "s" + std::to_string(s);
            sc.charge = species[5 * s];
            sc.mass = species[5 * s + 1];
            sc.weight = species[5 * s + 2];
            sc.particles = static_cast<std::size_t>(species[5 * s + 3]);
            sc.thermal_velocity = species[5 * s + 4];
            cfg.species.push_back(sc);
        }
        pic::Simulation3D sim(cfg);
        sim.initialize();
        auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; ++s) sim.step();
        auto t1 = std::chrono::steady_clock::now();
        *seconds_per_step = std::chrono::duration<double>(t1 - t0).count() / steps;
        int64_t live = 0;
        for (const auto& sp : sim.species()) live += static_cast<int64_t>(sp.live_count());
        *live_particles = live;
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Runs the reference's own 1D Simulation::run() (collisionless) and lets it
// write its outputs (scalars.csv, fields_<step>.csv, boundary_losses.csv,
// power_transfer.csv, ...) into output_dir.  electrodes: left then right
// {offset, amplitude, frequency, phase}.  species: nspecies rows of
// (charge, mass, weight, particles, drift_velocity, thermal_velocity,
// init_x_min, init_x_max); names: nspecies NUL-terminated strings.  An
// empty aps_path loads the species the reference's way from seed.
int ref_run1d(const char* output_dir, const char* aps_path, int64_t nx, double length, double dt,
              int64_t steps, int64_t output_interval, int dirichlet, const double* electrodes,
              int velocity_dimensions, uint64_t seed, int nspecies, const char* const* names,
              const double* species, char* err, int errlen) {
    try {
        pic::Config cfg;
        cfg.velocity_dimensions = static_cast<std::size_t>(velocity_dimensions);
        cfg.nx = static_cast<std::size_t>(nx);
        cfg.length = length;
        cfg.dt = dt;
        cfg.steps = static_cast<std::size_t>(steps);
        cfg.output_interval = static_cast<std::size_t>(output_interval);
        cfg.output_dir = output_dir;
        cfg.boundary = dirichlet ? pic::Boundary::Dirichlet : pic::Boundary::Periodic;
        cfg.phi_left = electrodes[0];
        cfg.phi_left_drive = {electrodes[1], electrodes[2], electrodes[3]};
        cfg.phi_right = electrodes[4];
        cfg.phi_right_drive = {electrodes[5], electrodes[6], electrodes[7]};
        cfg.seed = static_cast<unsigned>(seed);
        if (aps_path && *aps_path) cfg.initial_state_path = aps_path;
        for (int k = 0; k < nspecies; ++k) {
            const double* r = species + 8 * k;
            pic::SpeciesConfig sc;
            sc.name = names[k];
            sc.charge = r[0];
            sc.mass = r[1];
            sc.weight = r[2];
            sc.particles = static_cast<std::size_t>(r[3]);
            sc.drift_velocity = r[4];
            sc.thermal_velocity = r[5];
            sc.init_x_min = r[6];
            sc.init_x_max = r[7];
            cfg.species.push_back(sc);
        }
        pic::Simulation sim(cfg);
        sim.run();
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// The reference's own Simulation3D::run() on a periodic mesh (species
// loaded the reference's way from seed; rows of (charge, mass, weight,
// particles, thermal_velocity)), with an optional tabulated B along
// b_axis, writing scalars.csv into output_dir.
int ref_run3d(const char* output_dir, const int64_t* n, const double* L, double dt,
              int64_t steps, int64_t output_interval, uint64_t seed, int b_axis, int64_t brows,
              const double* bcoords, const double* bvalues, int nspecies,
              const char* const* names, const double* species, char* err, int errlen) {
    try {
        pic::Simulation3DConfig cfg;
        cfg.nx = static_cast<std::size_t>(n[0]);
        cfg.ny = static_cast<std::size_t>(n[1]);
        cfg.nz = static_cast<std::size_t>(n[2]);
        cfg.length_x = L[0];
        cfg.length_y = L[1];
        cfg.length_z = L[2];
        cfg.dt = dt;
        cfg.steps = static_cast<std::size_t>(steps);
        cfg.output_interval = static_cast<std::size_t>(output_interval);
        cfg.output_dir = output_dir;
        cfg.seed = static_cast<unsigned>(seed);
        cfg.boundary = pic::Boundary::Periodic;
        if (brows > 0) {
            std::vector<double> c(bcoords, bcoords + brows);
            std::vector<pic::Vec3> v(static_cast<std::size_t>(brows));
            for (int64_t i = 0; i < brows; ++i)
                v[i] = {bvalues[3 * i], bvalues[3 * i + 1], bvalues[3 * i + 2]};
            const auto axis = b_axis == 0 ? pic::CoordinateAxis::X
                              : b_axis == 1 ? pic::CoordinateAxis::Y : pic::CoordinateAxis::Z;
            cfg.magnetic_field_profile = pic::TabulatedVectorField1D(axis, c, v);
        }
        for (int k = 0; k < nspecies; ++k) {
            const double* r = species + 5 * k;
            pic::Species3DConfig sc;
            sc.name = names[k];
            sc.charge = r[0];
            sc.mass = r[1];
            sc.weight = r[2];
            sc.particles = static_cast<std::size_t>(r[3]);
            sc.thermal_velocity = r[4];
            cfg.species.push_back(sc);
        }
        pic::Simulation3D sim(cfg);
        sim.run();
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// The reference's own Simulation2D::run() on a periodic mesh with a
// uniform B (any nonzero component selects its 2D3V Boris push), species
// loaded from seed; rows of (charge, mass, weight, particles,
// thermal_velocity); writes scalars.csv into output_dir.
int ref_run2d(const char* output_dir, const int64_t* n, const double* L, double depth, double dt,
              int64_t steps, int64_t output_interval, uint64_t seed, const double* B,
              int nspecies, const char* const* names, const double* species, char* err,
              int errlen) {
    try {
        pic::Simulation2DConfig cfg;
        cfg.nx = static_cast<std::size_t>(n[0]);
        cfg.ny = static_cast<std::size_t>(n[1]);
        cfg.length_x = L[0];
        cfg.length_y = L[1];
        cfg.out_of_plane_depth = depth;
        cfg.dt = dt;
        cfg.steps = static_cast<std::size_t>(steps);
        cfg.output_interval = static_cast<std::size_t>(output_interval);
        cfg.output_dir = output_dir;
        cfg.seed = static_cast<unsigned>(seed);
        cfg.boundary = pic::Boundary::Periodic;
        cfg.magnetic_field_x = B[0];
        cfg.magnetic_field_y = B[1];
        cfg.magnetic_field_z = B[2];
        for (int k = 0; k < nspecies; ++k) {
            const double* r = species + 5 * k;
            pic::Species2DConfig sc;
            sc.name = names[k];
            sc.charge = r[0];
            sc.mass = r[1];
            sc.weight = r[2];
            sc.particles = static_cast<std::size_t>(r[3]);
            sc.thermal_velocity = r[4];
            cfg.species.push_back(sc);
        }
        pic::Simulation2D sim(cfg);
        sim.run();
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}


// The reference's own 1D Simulation::run() on a case file (the format of
// tests/golden/run1d/<case>/case.txt, read the same way by the host test
// program's --run1d mode): mesh, drive, species lines, and optionally
// "units si", "max_particles N" and named null-collision models
//   model <name> <target species> <neutral_density> <max_frequency>
//         <neutral_mass> <max_candidates> <neutral_temperature K>
//   channel <name> <elastic|excitation|ionization> <threshold J>
//           <table file, energies J and m^2> <secondary|-> <ion|->
// (channels belong to the model above them).  seed >= 0 replaces the case's
// seed.  After the run, final_<species>.f64 in output_dir holds the live
// particles as (x, v, velocity_y, velocity_z) doubles.
int ref_run1d_case(const char* case_path, const char* output_dir, int64_t seed, char* err,
                   int errlen) {
    try {
        const std::filesystem::path path(case_path);
        std::ifstream in(path);
        if (!in) throw std::runtime_error(std::string("cannot open ") + case_path);
        pic::Config cfg;
        cfg.output_dir = output_dir;
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string key;
            if (!(ls >> key)) continue;
            if (key == "nx") ls >> cfg.nx;
            else if (key == "length") ls >> cfg.length;
            else if (key == "dt") ls >> cfg.dt;
            else if (key == "steps") ls >> cfg.steps;
            else if (key == "output_interval") ls >> cfg.output_interval;
            else if (key == "velocity_dimensions") ls >> cfg.velocity_dimensions;
            else if (key == "seed") ls >> cfg.seed;
            else if (key == "precision") { int p; ls >> p; }
            else if (key == "max_particles") ls >> cfg.max_particles_per_species;
            else if (key == "units") {
                std::string u;
                ls >> u;
                cfg.units.system = u == "si" ? pic::UnitSystem::SI : pic::UnitSystem::Normalized;
            } else if (key == "dirichlet") {
                int v = 0;
                ls >> v;
                cfg.boundary = v ? pic::Boundary::Dirichlet : pic::Boundary::Periodic;
            } else if (key == "electrodes") {
                ls >> cfg.phi_left >> cfg.phi_left_drive.amplitude >>
                    cfg.phi_left_drive.frequency >> cfg.phi_left_drive.phase >> cfg.phi_right >>
                    cfg.phi_right_drive.amplitude >> cfg.phi_right_drive.frequency >>
                    cfg.phi_right_drive.phase;
            } else if (key == "initial_state") {
                std::string f;
                ls >> f;
                cfg.initial_state_path = path.parent_path() / f;
            } else if (key == "species") {
                pic::SpeciesConfig sc;
                ls >> sc.name >> sc.charge >> sc.mass >> sc.weight >> sc.particles >>
                    sc.drift_velocity >> sc.thermal_velocity >> sc.init_x_min >> sc.init_x_max;
                cfg.species.push_back(sc);
            } else if (key == "model") {
                pic::NamedCollisionConfig m;
                std::size_t cand = 64;
                ls >> m.name >> m.config.species >> m.config.neutral_density >>
                    m.config.max_frequency >> m.config.neutral_mass >> cand >>
                    m.config.neutral_temperature;
                m.config.enabled = true;
                m.config.model = pic::CollisionModelKind::NullCollision;
                m.config.max_candidates_per_particle = cand;
                m.config.gas_data_units = pic::UnitSystem::SI;
                cfg.collision_models.push_back(m);
            } else if (key == "channel") {
                if (cfg.collision_models.empty())
                    throw std::runtime_error("channel line before any model line");
                pic::CollisionChannelConfig ch;
                std::string kind, table, sec, ion;
                ls >> ch.name >> kind >> ch.threshold_energy >> table >> sec >> ion;
                ch.process = kind == "elastic"      ? pic::CollisionProcessKind::Elastic
                             : kind == "excitation" ? pic::CollisionProcessKind::Excitation
                             : kind == "ionization" ? pic::CollisionProcessKind::Ionization
                             : kind == "charge_exchange"
                                 ? pic::CollisionProcessKind::ChargeExchange
                                 : throw std::runtime_error("unknown channel kind " + kind);
                ch.cross_section_file = path.parent_path() / table;
                if (sec != "-") ch.secondary_species = sec;
                if (ion != "-") ch.ion_species = ion;
                cfg.collision_models.back().config.channels.push_back(ch);
            } else {
                throw std::runtime_error("unknown case key " + key);
            }
            if (ls.fail()) throw std::runtime_error("malformed case line: " + line);
        }
        if (seed >= 0) cfg.seed = static_cast<unsigned>(seed);
        pic::Simulation sim(cfg);
        sim.run();
        for (const auto& sp : sim.species()) {
            std::ofstream out(std::filesystem::path(output_dir) /
                                  ("final_" + sp.config().name + ".f64"),
                              std::ios::binary);
            for (const auto& p : sp.particles()) {
                if (!p.alive) continue;
                const double row[4] = {p.x, p.v, p.velocity_y, p.velocity_z};
                out.write(reinterpret_cast<const char*>(row), sizeof row);
            }
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

}  // extern "C"
