// particle_state.cpp - .aps version 1 reader, writer and signature
// (ParticleState.cpp:1-629 of the reference, restated).
#include "particle_state.hpp"

#include <bit>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <limits>
#include <set>
#include <stdexcept>

namespace pic::gpu {

namespace {

constexpr const char* kMagic = "AuroraPIC-particle-state-v1";
constexpr std::uint64_t kFnvOffset = 14695981039346656037ULL;
constexpr std::uint64_t kFnvPrime = 1099511628211ULL;

// FNV-1a over the little-endian bytes of a 64-bit value
void mix_u64(std::uint64_t& h, std::uint64_t v) {
    for (unsigned b = 0; b < 8; ++b) {
        h ^= static_cast<unsigned char>(v >> (8 * b));
        h *= kFnvPrime;
    }
}

// length first, then the bytes
void mix_string(std::uint64_t& h, const std::string& s) {
    mix_u64(h, s.size());
    for (const unsigned char c : s) {
        h ^= c;
        h *= kFnvPrime;
    }
}

void mix_double(std::uint64_t& h, double v) { mix_u64(h, std::bit_cast<std::uint64_t>(v)); }

void mix_record(std::uint64_t& h, const ExternalParticleRecord& r) {
    mix_double(h, r.position.x);
    mix_double(h, r.position.y);
    mix_double(h, r.position.z);
    mix_double(h, r.velocity.x);
    mix_double(h, r.velocity.y);
    mix_double(h, r.velocity.z);
}

bool all_finite(const ExternalParticleRecord& r) {
    return std::isfinite(r.position.x) && std::isfinite(r.position.y) &&
           std::isfinite(r.position.z) && std::isfinite(r.velocity.x) &&
           std::isfinite(r.velocity.y) && std::isfinite(r.velocity.z);
}

void expect_key(std::istream& in, const std::string& key, const std::filesystem::path& path) {
    std::string got;
    in >> got;
    if (!in || got != key)
        throw std::runtime_error("external particle state '" + path.string() +
                                 "' expected key '" + key + "'");
}

UnitSystem units_from(const std::string& v, const std::filesystem::path& path) {
    if (v == "normalized") return UnitSystem::Normalized;
    if (v == "si") return UnitSystem::SI;
    throw std::runtime_error("external particle state '" + path.string() +
                             "' has invalid units '" + v + "'");
}

// ParticleState.cpp validate_structure
void check_structure(const ExternalParticleState& s, const std::string& context) {
    if (s.version != 1 || s.spatial_dimension == 0 || s.spatial_dimension > 3 ||
        s.particle_count == 0)
        throw std::invalid_argument(context + " has invalid particle-state metadata");
    std::size_t total = 0;
    for (const auto& [name, records] : s.species) {
        if (name.empty() || records.empty())
            throw std::invalid_argument(context + " has an empty species name or population");
        if (records.size() > std::numeric_limits<std::size_t>::max() - total)
            throw std::overflow_error(context + " particle count overflow");
        total += records.size();
        for (const auto& r : records) {
            const bool inactive_set =
                (s.spatial_dimension < 3 && r.position.z != 0.0) ||
                (s.spatial_dimension == 1 &&
                 (r.position.y != 0.0 || r.velocity.y != 0.0 || r.velocity.z != 0.0));
            if (!all_finite(r) || inactive_set)
                throw std::invalid_argument(
                    context + " has invalid or active out-of-dimension components");
        }
    }
    if (total != s.particle_count)
        throw std::invalid_argument(context + " particle count does not match its records");
}

}  // namespace

std::string to_string(UnitSystem units) {
    return units == UnitSystem::SI ? "si" : "normalized";
}

ExternalParticleState load_external_particle_state(const std::filesystem::path& path,
                                                   std::size_t max_particles) {
    if (max_particles == 0)
        throw std::invalid_argument("external particle-state limit must be positive");
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open external particle state: " + path.string());
    std::string magic;
    in >> magic;
    if (magic != kMagic)
        throw std::runtime_error("invalid external particle-state magic in: " + path.string());
    ExternalParticleState s;
    expect_key(in, "dimension", path);
    in >> s.spatial_dimension;
    if (!in || s.spatial_dimension == 0 || s.spatial_dimension > 3)
        throw std::runtime_error("external particle state has invalid dimension");
    expect_key(in, "units", path);
    std::string units;
    in >> units;
    s.unit_system = units_from(units, path);
    expect_key(in, "weighting", path);
    std::string weighting;
    in >> weighting;
    if (weighting != "species_constant")
        throw std::runtime_error(
            "external particle state supports only weighting species_constant");
    expect_key(in, "velocity_staggering", path);
    std::string staggering;
    in >> staggering;
    if (staggering != "time_centered")
        throw std::runtime_error(
            "external particle state supports only time_centered velocities");
    expect_key(in, "particle_count", path);
    in >> s.particle_count;
    if (!in || s.particle_count == 0 || s.particle_count > max_particles)
        throw std::runtime_error(
            "external particle-state count is zero, invalid, or exceeds the configured limit");
    expect_key(in, "records", path);
    for (std::size_t i = 0; i < s.particle_count; ++i) {
        expect_key(in, "particle", path);
        std::string name;
        ExternalParticleRecord r;
        in >> name >> r.position.x >> r.position.y >> r.position.z >> r.velocity.x >>
            r.velocity.y >> r.velocity.z;
        if (!in || name.empty() || !all_finite(r))
            throw std::runtime_error("invalid external particle record " + std::to_string(i) +
                                     " in: " + path.string());
        if (s.spatial_dimension < 3 && r.position.z != 0.0)
            throw std::runtime_error("external particle state has nonzero inactive z position");
        if (s.spatial_dimension == 1 &&
            (r.position.y != 0.0 || r.velocity.y != 0.0 || r.velocity.z != 0.0))
            throw std::runtime_error(
                "1D1V external particle state has nonzero inactive components");
        s.species[name].push_back(r);
    }
    expect_key(in, "end", path);
    std::string trailing;
    if (in >> trailing)
        throw std::runtime_error("external particle state contains trailing data: " +
                                 path.string());
    s.signature = external_particle_state_signature(s);
    return s;
}

std::uint64_t external_particle_state_signature(const ExternalParticleState& s) {
    check_structure(s, "external particle state");
    std::uint64_t h = kFnvOffset;
    mix_string(h, kMagic);
    mix_u64(h, s.version);
    mix_u64(h, s.spatial_dimension);
    mix_string(h, to_string(s.unit_system));
    mix_u64(h, s.particle_count);
    mix_u64(h, s.species.size());
    for (const auto& [name, records] : s.species) {
        mix_string(h, name);
        mix_u64(h, records.size());
        for (const auto& r : records) mix_record(h, r);
    }
    return h;
}

void validate_external_particle_state(const ExternalParticleState& s,
                                      std::size_t spatial_dimension, UnitSystem unit_system,
                                      const std::vector<ExternalSpeciesExpectation>& expected,
                                      const std::string& context) {
    check_structure(s, context);
    if (s.version != 1)
        throw std::invalid_argument(context + " external particle-state version is unsupported");
    if (s.spatial_dimension != spatial_dimension)
        throw std::invalid_argument(context +
                                    " external particle-state dimension does not match");
    if (s.unit_system != unit_system)
        throw std::invalid_argument(context +
                                    " external particle-state unit system does not match");
    std::set<std::string> names;
    std::size_t total = 0;
    for (const auto& e : expected) {
        if (e.name.empty() || !names.insert(e.name).second)
            throw std::invalid_argument(context +
                                        " requires unique non-empty configured species names");
        if (e.particle_count > std::numeric_limits<std::size_t>::max() - total)
            throw std::overflow_error(context + " configured particle count overflow");
        total += e.particle_count;
        const auto it = s.species.find(e.name);
        if (it == s.species.end())
            throw std::invalid_argument(context + " external particle state is missing species '" +
                                        e.name + "'");
        if (it->second.size() != e.particle_count)
            throw std::invalid_argument(context + " external particle count for species '" +
                                        e.name + "' does not match config");
    }
    for (const auto& [name, records] : s.species) {
        (void)records;
        if (!names.contains(name))
            throw std::invalid_argument(context +
                                        " external particle state contains unknown species '" +
                                        name + "'");
    }
    if (s.particle_count != total)
        throw std::invalid_argument(context +
                                    " external total particle count does not match config");
}

ExternalParticleState load_validated_external_particle_state(
    const std::filesystem::path& path, std::size_t spatial_dimension, UnitSystem unit_system,
    const std::vector<ExternalSpeciesExpectation>& expected, const std::string& context,
    std::optional<std::uint64_t> expected_signature) {
    std::size_t total = 0;
    for (const auto& e : expected) {
        if (e.particle_count > std::numeric_limits<std::size_t>::max() - total)
            throw std::overflow_error(context + " configured particle count overflow");
        total += e.particle_count;
    }
    if (total == 0)
        throw std::invalid_argument(context + " external state requires configured particles");
    auto s = load_external_particle_state(path, total);
    validate_external_particle_state(s, spatial_dimension, unit_system, expected, context);
    if (expected_signature && s.signature != *expected_signature)
        throw std::invalid_argument(context +
                                    " external particle-state signature mismatch: expected " +
                                    std::to_string(*expected_signature) + ", got " +
                                    std::to_string(s.signature));
    return s;
}

void write_external_particle_state(const std::filesystem::path& path,
                                   const ExternalParticleState& s, bool overwrite) {
    check_structure(s, "external particle state writer");
    if (!overwrite && std::filesystem::exists(path))
        throw std::runtime_error("external particle-state output already exists: " +
                                 path.string());
    if (!path.parent_path().empty()) std::filesystem::create_directories(path.parent_path());
    std::ofstream out(path, std::ios::trunc);
    if (!out) throw std::runtime_error("cannot write external particle state: " + path.string());
    out << kMagic << '\n'
        << "dimension " << s.spatial_dimension << '\n'
        << "units " << to_string(s.unit_system) << '\n'
        << "weighting species_constant\n"
        << "velocity_staggering time_centered\n"
        << "particle_count " << s.particle_count << '\n'
        << "records\n"
        << std::setprecision(17);
    for (const auto& [name, records] : s.species)
        for (const auto& r : records)
            out << "particle " << name << ' ' << r.position.x << ' ' << r.position.y << ' '
                << r.position.z << ' ' << r.velocity.x << ' ' << r.velocity.y << ' '
                << r.velocity.z << '\n';
    out << "end\n";
    if (!out)
        throw std::runtime_error("failed while writing external particle state: " +
                                 path.string());
}

void write_external_particle_state_metadata(const std::filesystem::path& path,
                                            const std::filesystem::path& source_path,
                                            const ExternalParticleStateMetadata& m,
                                            std::optional<std::uint64_t> expected_signature) {
    if (m.version != 1 || m.spatial_dimension == 0 || m.spatial_dimension > 3 ||
        m.particle_count == 0)
        throw std::invalid_argument("external particle-state metadata is invalid");
    if (!path.parent_path().empty()) std::filesystem::create_directories(path.parent_path());
    std::ofstream out(path);
    if (!out)
        throw std::runtime_error("cannot write external particle-state metadata: " +
                                 path.string());
    out << "format aurorapic_particle_state\n"
        << "version " << m.version << '\n'
        << "source_path " << std::quoted(source_path.string()) << '\n'
        << "dimension " << m.spatial_dimension << '\n'
        << "units " << to_string(m.unit_system) << '\n'
        << "particle_count " << m.particle_count << '\n'
        << "signature " << m.signature << '\n'
        << "expected_signature ";
    if (expected_signature)
        out << *expected_signature;
    else
        out << "none";
    out << '\n';
    if (!out)
        throw std::runtime_error("failed while writing external particle-state metadata: " +
                                 path.string());
}

}  // namespace pic::gpu
