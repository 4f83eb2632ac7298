// diagnostics.cpp - writers of the reference's run outputs.
#include "diagnostics.hpp"

#include <cmath>
#include <iomanip>
#include <stdexcept>
#include <utility>

namespace pic::gpu {

namespace {

void check_permittivity(double eps, const char* what) {
    if (!std::isfinite(eps) || !(eps > 0.0)) throw std::invalid_argument(what);
}

std::ofstream open_scalars(const std::filesystem::path& dir, const char* what) {
    std::filesystem::create_directories(dir);
    std::ofstream f(dir / "scalars.csv");
    if (!f) throw std::runtime_error(what);
    return f;
}

// the leading columns every scalars.csv shares
void write_common(std::ostream& out, const DiagnosticSample& s) {
    out << s.step << ',' << std::setprecision(17) << s.time << ',' << s.kinetic_energy << ','
        << s.field_energy << ',' << s.total_energy << ',' << s.charge_l1 << ','
        << s.live_particles;
}

void write_live_by_species(std::ostream& out, const DiagnosticSample& s) {
    for (const auto live : s.live_particles_by_species) out << ',' << live;
    out << '\n';
    out.flush();
}

void live_headers(std::ostream& out, const std::vector<std::string>& names) {
    for (const auto& name : names) out << ",live_particles_" << name;
    out << '\n';
}

const char* energy_suffix(bool si) { return si ? "_J_m-2" : "_normalized"; }

}  // namespace

std::string csv_cell(const std::string& value) {
    if (value.find_first_of(",\"\r\n") == std::string::npos) return value;
    std::string out = "\"";
    for (const char ch : value) {
        out.push_back(ch);
        if (ch == '"') out.push_back('"');
    }
    out.push_back('"');
    return out;
}

// ------------------------------------------------------------------ 1D
Diagnostics::Diagnostics(std::filesystem::path output_dir, std::vector<std::string> species_names,
                         double permittivity)
    : output_dir_(std::move(output_dir)), species_names_(std::move(species_names)) {
    check_permittivity(permittivity, "diagnostic permittivity must be positive and finite");
    scalar_file_ = open_scalars(output_dir_, "cannot open diagnostics output");
}

void Diagnostics::write_header() {
    scalar_file_ << "step,time,kinetic_energy,field_energy,total_energy,charge_l1,"
                    "live_particles,phi_left,phi_right";
    live_headers(scalar_file_, species_names_);
}

DiagnosticSample Diagnostics::sample(const Simulation& sim) {
    history_.push_back(sim.sample());
    return history_.back();
}

void Diagnostics::write_sample(const DiagnosticSample& s) {
    write_common(scalar_file_, s);
    scalar_file_ << ',' << s.phi_left << ',' << s.phi_right;
    write_live_by_species(scalar_file_, s);
}

void Diagnostics::write_fields(std::size_t step, const Grid& grid) const {
    std::ofstream out(output_dir_ / ("fields_" + std::to_string(step) + ".csv"));
    out << "x,rho,phi,E\n" << std::setprecision(17);
    for (std::size_t i = 0; i < grid.nx(); ++i)
        out << grid.node_x(i) << ',' << grid.rho()[i] << ',' << grid.phi()[i] << ','
            << grid.electric()[i] << '\n';
}

// ------------------------------------------------------------------ 2D
Diagnostics2D::Diagnostics2D(std::filesystem::path output_dir,
                             std::vector<std::string> species_names, double permittivity,
                             double out_of_plane_depth)
    : output_dir_(std::move(output_dir)), species_names_(std::move(species_names)) {
    check_permittivity(permittivity, "2D diagnostic permittivity must be positive and finite");
    if (!std::isfinite(out_of_plane_depth) || !(out_of_plane_depth > 0.0))
        throw std::invalid_argument("2D diagnostic depth must be positive and finite");
    scalar_file_ = open_scalars(output_dir_, "cannot open 2D diagnostics output");
}

void Diagnostics2D::write_header() {
    scalar_file_ << "step,time,kinetic_energy,field_energy,total_energy,charge_l1,live_particles"
                 << ",absorbed_left,absorbed_right,absorbed_bottom,absorbed_top";
    live_headers(scalar_file_, species_names_);
}

DiagnosticSample Diagnostics2D::sample(const Simulation2D& sim) {
    history_.push_back(sim.sample());
    return history_.back();
}

void Diagnostics2D::write_sample(const DiagnosticSample& s) {
    write_common(scalar_file_, s);
    for (int f = 0; f < 4; ++f) scalar_file_ << ',' << s.absorbed[f];
    write_live_by_species(scalar_file_, s);
}

// ------------------------------------------------------------------ 3D
Diagnostics3D::Diagnostics3D(std::filesystem::path output_dir,
                             std::vector<std::string> species_names, double permittivity)
    : output_dir_(std::move(output_dir)), species_names_(std::move(species_names)) {
    check_permittivity(permittivity, "3D diagnostic permittivity must be positive and finite");
    scalar_file_ = open_scalars(output_dir_, "cannot open 3D diagnostics output");
}

void Diagnostics3D::write_header() {
    scalar_file_ << "step,time,kinetic_energy,field_energy,total_energy,charge_l1,live_particles"
                 << ",absorbed_left,absorbed_right,absorbed_bottom,absorbed_top,absorbed_back,"
                    "absorbed_front";
    live_headers(scalar_file_, species_names_);
}

DiagnosticSample Diagnostics3D::sample(const Simulation3D& sim) {
    history_.push_back(sim.sample());
    return history_.back();
}

void Diagnostics3D::write_sample(const DiagnosticSample& s) {
    write_common(scalar_file_, s);
    for (int f = 0; f < 6; ++f) scalar_file_ << ',' << s.absorbed[f];
    write_live_by_species(scalar_file_, s);
}

// ------------------------------------------------------- 1D ledgers
void write_collision_header(std::ostream& out, const std::vector<std::string>& channel_names,
                            bool si) {
    const char* suffix = energy_suffix(si);
    out << "step,time,candidates,null_collisions";
    for (const auto& n : channel_names) out << ",collisions_" << n;
    for (const auto& n : channel_names) out << ",tracked_kinetic_energy_change_" << n << suffix;
    out << ",cumulative_candidates,cumulative_null_collisions";
    for (const auto& n : channel_names) out << ",cumulative_collisions_" << n;
    for (const auto& n : channel_names)
        out << ",cumulative_tracked_kinetic_energy_change_" << n << suffix;
    out << '\n';
}

void write_collision_sample(std::ostream& out, std::size_t step, double time,
                            const CollisionDiagnostics& interval,
                            const CollisionDiagnostics& totals) {
    auto body = [&](const CollisionDiagnostics& d) {
        out << d.candidates << ',' << d.null_collisions;
        for (const auto n : d.channel_collisions) out << ',' << n;
        for (const auto e : d.channel_energy_change) out << ',' << e;
    };
    out << step << ',' << std::setprecision(17) << time << ',';
    body(interval);
    out << ',';
    body(totals);
    out << '\n';
    out.flush();
}

void write_boundary_loss_header(std::ostream& out, const std::vector<std::string>& species_names,
                                bool si) {
    const std::string charge_unit = si ? "_C_m-2" : "_normalized";
    const std::string energy_unit = energy_suffix(si);
    out << "step,time,counter_origin_step";
    for (const auto& name : species_names) {
        out << ',' << csv_cell("absorbed_left_count_" + name) << ','
            << csv_cell("absorbed_right_count_" + name) << ','
            << csv_cell("absorbed_left_charge_" + name + charge_unit) << ','
            << csv_cell("absorbed_right_charge_" + name + charge_unit) << ','
            << csv_cell("absorbed_left_kinetic_energy_" + name + energy_unit) << ','
            << csv_cell("absorbed_right_kinetic_energy_" + name + energy_unit);
    }
    out << '\n';
}

void write_boundary_loss_sample(std::ostream& out, std::size_t step, double time,
                                std::size_t counter_origin_step,
                                const std::vector<double>& charge_per_macro,
                                const std::vector<BoundaryLoss1D>& losses) {
    if (charge_per_macro.size() != losses.size())
        throw std::logic_error("1D boundary-loss diagnostics do not match species");
    out << step << ',' << std::setprecision(17) << time << ',' << counter_origin_step;
    for (std::size_t k = 0; k < losses.size(); ++k) {
        const BoundaryLoss1D& l = losses[k];
        out << ',' << l.absorbed_left << ',' << l.absorbed_right << ','
            << charge_per_macro[k] * static_cast<double>(l.absorbed_left) << ','
            << charge_per_macro[k] * static_cast<double>(l.absorbed_right) << ','
            << l.kinetic_energy_left << ',' << l.kinetic_energy_right;
    }
    out << '\n';
    out.flush();
}

void write_power_transfer_header(std::ostream& out,
                                 const std::vector<std::string>& species_names, bool si) {
    out << "step,time,counter_origin_step";
    for (const auto& name : species_names)
        out << ',' << csv_cell("electric_work_" + name + energy_suffix(si));
    out << '\n';
}

void write_power_transfer_sample(std::ostream& out, std::size_t step, double time,
                                 std::size_t counter_origin_step,
                                 const std::vector<SpeciesPower1D>& power) {
    out << step << ',' << std::setprecision(17) << time << ',' << counter_origin_step;
    for (const auto& p : power) out << ',' << p.electric_work;
    out << '\n';
    out.flush();
}

}  // namespace pic::gpu
