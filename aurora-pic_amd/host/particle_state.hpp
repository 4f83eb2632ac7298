// particle_state.hpp - the reference's external particle-state interchange
// format (.aps version 1, ParticleState.hpp / ParticleState.cpp:1-629;
// docs/external-particle-state.md): a time-zero particle population with
// time-centred velocities, its canonical 64-bit FNV-1a signature, the
// reader (whole-state and validated forms), the writer and the metadata
// file.  Simulation*::initialize() of the GPU host classes ingest it the way
// the reference simulations do.
//
// Host-side I/O only (no device calls): the same functions, argument
// meanings, exception classes and messages as the reference, so files and
// signatures interchange with it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <map>
#include <optional>
#include <string>
#include <vector>

#include "aurora_gpu.hpp"

namespace pic::gpu {

std::string to_string(UnitSystem units);

struct ExternalParticleRecord {
    Vec3 position{};
    Vec3 velocity{};      // time-centred
};

// species name -> records in file order; std::map keeps the canonical
// (sorted-name) order the signature uses
struct ExternalParticleState {
    std::size_t version{1};
    std::size_t spatial_dimension{0};
    UnitSystem unit_system{UnitSystem::Normalized};
    std::map<std::string, std::vector<ExternalParticleRecord>> species;
    std::size_t particle_count{0};
    std::uint64_t signature{0};
};

struct ExternalParticleStateMetadata {
    std::size_t version{1};
    std::size_t spatial_dimension{0};
    UnitSystem unit_system{UnitSystem::Normalized};
    std::size_t particle_count{0};
    std::uint64_t signature{0};
};

struct ExternalSpeciesExpectation {
    std::string name;
    std::size_t particle_count{0};
};

// Reads a version-1 file (at most max_particles records); the signature of
// the loaded state is filled in.
ExternalParticleState load_external_particle_state(const std::filesystem::path& path,
                                                   std::size_t max_particles);

// The canonical FNV-1a signature: magic, version, dimension, units, count,
// then per species in name order its name, population and every binary64
// position and velocity component.
std::uint64_t external_particle_state_signature(const ExternalParticleState& state);

// Dimension, units, species names and per-species counts against a run.
void validate_external_particle_state(const ExternalParticleState& state,
                                      std::size_t spatial_dimension, UnitSystem unit_system,
                                      const std::vector<ExternalSpeciesExpectation>& expected,
                                      const std::string& context);

// Load + validate + optional pinned signature.
ExternalParticleState load_validated_external_particle_state(
    const std::filesystem::path& path, std::size_t spatial_dimension, UnitSystem unit_system,
    const std::vector<ExternalSpeciesExpectation>& expected, const std::string& context,
    std::optional<std::uint64_t> expected_signature = std::nullopt);

// Sorted species order, binary64 round-trip precision; refuses to replace an
// existing file unless overwrite.
void write_external_particle_state(const std::filesystem::path& path,
                                   const ExternalParticleState& state, bool overwrite = false);

void write_external_particle_state_metadata(const std::filesystem::path& path,
                                            const std::filesystem::path& source_path,
                                            const ExternalParticleStateMetadata& metadata,
                                            std::optional<std::uint64_t> expected_signature);

}  // namespace pic::gpu
