// diagnostics.hpp - the reference's run outputs for device simulations.
//
// Diagnostics / Diagnostics2D / Diagnostics3D write scalars.csv with the
// columns, number format (max_digits10) and validation messages of
// Diagnostics.cpp:10-230; the 1D run also writes fields_<step>.csv and the
// collision, boundary-loss and power-transfer ledgers of
// Simulation.cpp:154-297.  Samples come from the device
// (Simulation*::sample()); these classes only format and write them.
#pragma once

#include <cstddef>
#include <filesystem>
#include <fstream>
#include <ostream>
#include <string>
#include <vector>

#include "aurora_gpu.hpp"

namespace pic::gpu {

// CSV quoting of Simulation.cpp:299-310: cells holding a comma, quote or
// line break are quoted, quotes doubled.
std::string csv_cell(const std::string& value);

// scalars.csv of a 1D run (Diagnostics.cpp:10-72) and fields_<step>.csv.
class Diagnostics {
public:
    Diagnostics(std::filesystem::path output_dir, std::vector<std::string> species_names,
                double permittivity);
    void write_header();
    // sim.sample(), kept in history()
    DiagnosticSample sample(const Simulation& sim);
    void write_sample(const DiagnosticSample& s);
    void write_fields(std::size_t step, const Grid& grid) const;
    const std::vector<DiagnosticSample>& history() const { return history_; }

private:
    std::filesystem::path output_dir_;
    std::vector<std::string> species_names_;
    std::ofstream scalar_file_;
    std::vector<DiagnosticSample> history_;
};

// scalars.csv of a structured 2D run (Diagnostics.cpp:74-154).
class Diagnostics2D {
public:
    Diagnostics2D(std::filesystem::path output_dir, std::vector<std::string> species_names,
                  double permittivity, double out_of_plane_depth);
    void write_header();
    DiagnosticSample sample(const Simulation2D& sim);
    void write_sample(const DiagnosticSample& s);
    const std::vector<DiagnosticSample>& history() const { return history_; }

private:
    std::filesystem::path output_dir_;
    std::vector<std::string> species_names_;
    std::ofstream scalar_file_;
    std::vector<DiagnosticSample> history_;
};

// scalars.csv of a 3D run (Diagnostics.cpp:178-249).
class Diagnostics3D {
public:
    Diagnostics3D(std::filesystem::path output_dir, std::vector<std::string> species_names,
                  double permittivity);
    void write_header();
    DiagnosticSample sample(const Simulation3D& sim);
    void write_sample(const DiagnosticSample& s);
    const std::vector<DiagnosticSample>& history() const { return history_; }

private:
    std::filesystem::path output_dir_;
    std::vector<std::string> species_names_;
    std::ofstream scalar_file_;
    std::vector<DiagnosticSample> history_;
};

// collisions.csv (Simulation.cpp:154-204): the interval since the last
// sample, then the totals.
void write_collision_header(std::ostream& out, const std::vector<std::string>& channel_names,
                            bool si);
void write_collision_sample(std::ostream& out, std::size_t step, double time,
                            const CollisionDiagnostics& interval,
                            const CollisionDiagnostics& totals);

// boundary_losses.csv (Simulation.cpp:207-264); charge_per_macro[k] is
// species k's charge times its weight.
void write_boundary_loss_header(std::ostream& out, const std::vector<std::string>& species_names,
                                bool si);
void write_boundary_loss_sample(std::ostream& out, std::size_t step, double time,
                                std::size_t counter_origin_step,
                                const std::vector<double>& charge_per_macro,
                                const std::vector<BoundaryLoss1D>& losses);

// power_transfer.csv (Simulation.cpp:266-297).
void write_power_transfer_header(std::ostream& out,
                                 const std::vector<std::string>& species_names, bool si);
void write_power_transfer_sample(std::ostream& out, std::size_t step, double time,
                                 std::size_t counter_origin_step,
                                 const std::vector<SpeciesPower1D>& power);

}  // namespace pic::gpu
