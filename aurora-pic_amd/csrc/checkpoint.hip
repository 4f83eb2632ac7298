// checkpoint.hip - checkpoint and restart of a context's device state.
//
// The reference writes text restart files guarded by a physics signature
// (Simulation3D.cpp:367-469, Simulation.cpp:1812-2900, the collision model
// signature of Collision.cpp:365-577).  Here the state is a versioned binary
// dump, bit-exact by construction:
//   magic, version, physics signature, time, step;
//   per species: slots, next uid, the SoA arrays as stored (packed cell,
//   offsets, v^{n-1/2}, uid), the boundary counters and, in 1D, the wall
//   energy and electric-work ledger;
//   the MCC counters.
// The fields are not stored: loading deposits the restored particles and
// solves, which reproduces the fields of the saved step to the bit (integer
// deposit, the same solve).  The physics signature hashes what the caller
// configures (mesh, species, magnetic field, particle boundaries,
// electrodes, the MCC model and its seed); loading into a context configured
// differently fails.  Batched swarms and ensembles are not part of it.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "aurora_internal.hpp"

namespace aurora {

namespace {

constexpr char kMagic[16] = "AURORA-HIP-CKPT";
constexpr uint32_t kVersion = 2;   // 2: every null-collision model, velocity dimensions

struct Fnv {
    uint64_t h = 14695981039346656037ULL;
    void bytes(const void* p, size_t n) {
        const auto* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ULL;
        }
    }
    template <typename T>
    void pod(const T& v) { bytes(&v, sizeof v); }
    template <typename T>
    void vec(const std::vector<T>& v) {
        pod(static_cast<uint64_t>(v.size()));
        if (!v.empty()) bytes(v.data(), v.size() * sizeof(T));
    }
};

template <typename T>
std::vector<T> to_host(const DevBuf<T>& b, size_t n) {
    std::vector<T> v(n);
    if (n) AUR_HIP(hipMemcpy(v.data(), b.ptr, n * sizeof(T), hipMemcpyDeviceToHost));
    return v;
}

void write_raw(std::ofstream& out, const void* p, size_t n) {
    out.write(static_cast<const char*>(p), static_cast<std::streamsize>(n));
}

void read_raw(std::ifstream& in, void* p, size_t n, const std::string& path) {
    in.read(static_cast<char*>(p), static_cast<std::streamsize>(n));
    AUR_REQUIRE(in.gcount() == static_cast<std::streamsize>(n), AURORA_ERR_INVALID_ARGUMENT,
                "checkpoint file is truncated: " + path);
}

template <typename T>
void write_pod(std::ofstream& out, const T& v) { write_raw(out, &v, sizeof v); }

template <typename T>
T read_pod(std::ifstream& in, const std::string& path) {
    T v{};
    read_raw(in, &v, sizeof v, path);
    return v;
}

// device buffer <-> file, n elements of elem bytes
void dump(std::ofstream& out, const void* dev, size_t bytes) {
    std::vector<unsigned char> h(bytes);
    if (bytes) AUR_HIP(hipMemcpy(h.data(), dev, bytes, hipMemcpyDeviceToHost));
    write_raw(out, h.data(), bytes);
}

void restore(std::ifstream& in, void* dev, size_t bytes, const std::string& path) {
    std::vector<unsigned char> h(bytes);
    read_raw(in, h.data(), bytes, path);
    if (bytes) AUR_HIP(hipMemcpy(dev, h.data(), bytes, hipMemcpyHostToDevice));
}

}  // namespace

uint64_t physics_signature(const Context& c) {
    Fnv f;
    f.pod(c.dim);
    f.bytes(c.n, sizeof c.n);
    f.bytes(c.L, sizeof c.L);
    f.pod(c.bc);
    f.pod(c.eps);
    f.pod(c.depth);
    f.bytes(c.pb, sizeof c.pb);
    f.bytes(c.Buni, sizeof c.Buni);
    f.pod(c.has_profile);
    if (c.has_profile) {
        f.pod(c.baxis);
        f.vec(to_host(c.bcoords, static_cast<size_t>(c.brows)));
        f.vec(to_host(c.bvalues, static_cast<size_t>(3 * c.brows)));
    }
    f.bytes(c.left4, sizeof c.left4);
    f.bytes(c.right4, sizeof c.right4);
    f.pod(static_cast<uint64_t>(c.species.size()));
    for (const auto& sp : c.species) {
        f.pod(sp.charge);
        f.pod(sp.mass);
        f.pod(sp.weight);
        f.pod(sp.precision);
    }
    f.pod(c.velocity_dims);
    f.pod(c.mcc_models());
    for (int model = 0; model < c.mcc_models(); ++model) {
        const MccModelHost& m = c.mcc_model(model);
        f.pod(m.target);
        f.pod(m.nchan);
        for (int k = 0; k < m.nchan; ++k) {
            f.pod(m.kind[k]);
            f.pod(m.threshold[k]);
            f.pod(m.angular[k]);
            f.pod(m.frame[k]);
            f.vec(m.energy_host[k]);
            f.vec(m.sigma_host[k]);
            f.vec(m.mc_energy_host[k]);
            f.vec(m.mc_value_host[k]);
        }
        f.pod(m.neutral_density);
        f.pod(m.max_frequency);
        f.pod(m.neutral_mass);
        f.pod(m.neutral_vstd);
        f.pod(m.particle_mass);
        f.pod(m.max_candidates);
        f.pod(m.secondary_species);
        f.pod(m.ion_species);
        f.pod(m.seed);
    }
    return f.h;
}

void checkpoint_save(Context& c, const std::string& path) {
    AUR_REQUIRE(c.dim > 0, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
    AUR_HIP(hipStreamSynchronize(c.stream));
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    AUR_REQUIRE(out.good(), AURORA_ERR_INVALID_ARGUMENT, "cannot write checkpoint: " + path);
    write_raw(out, kMagic, sizeof kMagic);
    write_pod(out, kVersion);
    write_pod(out, physics_signature(c));
    write_pod(out, c.time);
    write_pod(out, c.step_count);
    write_pod(out, static_cast<uint64_t>(c.species.size()));
    for (const auto& sp : c.species) {
        const size_t n = static_cast<size_t>(sp.slots);
        write_pod(out, sp.slots);
        write_pod(out, sp.next_uid);
        dump(out, sp.cell.ptr, n * sizeof(int32_t));
        for (int a = 0; a < c.dim; ++a) dump(out, sp.off[a].ptr, n * sp.elem());
        for (int a = 0; a < 3; ++a) dump(out, sp.vel[a].ptr, n * sp.elem());
        dump(out, sp.uid.ptr, n * sizeof(uint32_t));
        dump(out, sp.absorbed.ptr, 7 * sizeof(unsigned long long));
        const bool has_ledger = sp.ledger.ptr != nullptr;
        write_pod(out, has_ledger);
        if (has_ledger) dump(out, sp.ledger.ptr, 3 * sizeof(double));
    }
    write_pod(out, c.mcc_models());
    for (int model = 0; model < c.mcc_models(); ++model) {
        dump(out, c.mcc_model(model).counts.ptr, (kMaxChannels + 3) * sizeof(unsigned long long));
        dump(out, c.mcc_model(model).energy_change.ptr, kMaxChannels * sizeof(double));
    }
    out.flush();
    AUR_REQUIRE(out.good(), AURORA_ERR_INVALID_ARGUMENT,
                "failed while writing checkpoint: " + path);
}

void checkpoint_load(Context& c, const std::string& path) {
    AUR_REQUIRE(c.dim > 0, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
    std::ifstream in(path, std::ios::binary);
    AUR_REQUIRE(in.good(), AURORA_ERR_INVALID_ARGUMENT, "cannot open checkpoint: " + path);
    char magic[16] = {};
    read_raw(in, magic, sizeof magic, path);
    AUR_REQUIRE(std::memcmp(magic, kMagic, sizeof kMagic) == 0, AURORA_ERR_INVALID_ARGUMENT,
                "not a checkpoint file: " + path);
    const auto version = read_pod<uint32_t>(in, path);
    AUR_REQUIRE(version == kVersion, AURORA_ERR_UNSUPPORTED,
                "unsupported checkpoint version " + std::to_string(version));
    const auto sig = read_pod<uint64_t>(in, path);
    const uint64_t mine = physics_signature(c);
    AUR_REQUIRE(sig == mine, AURORA_ERR_INVALID_ARGUMENT,
                "checkpoint physics signature mismatch: expected " + std::to_string(mine) +
                    ", got " + std::to_string(sig));
    AUR_HIP(hipStreamSynchronize(c.stream));
    const auto time = read_pod<double>(in, path);
    const auto step = read_pod<int64_t>(in, path);
    const auto nsp = read_pod<uint64_t>(in, path);
    AUR_REQUIRE(nsp == c.species.size(), AURORA_ERR_INVALID_ARGUMENT,
                "checkpoint species count does not match the context");
    for (auto& sp : c.species) {
        const auto slots = read_pod<int64_t>(in, path);
        const auto next_uid = read_pod<uint32_t>(in, path);
        AUR_REQUIRE(slots >= 0 && slots <= sp.capacity, AURORA_ERR_CAPACITY,
                    "checkpoint population exceeds the species capacity");
        const size_t n = static_cast<size_t>(slots);
        restore(in, sp.cell.ptr, n * sizeof(int32_t), path);
        for (int a = 0; a < c.dim; ++a) restore(in, sp.off[a].ptr, n * sp.elem(), path);
        for (int a = 0; a < 3; ++a) restore(in, sp.vel[a].ptr, n * sp.elem(), path);
        restore(in, sp.uid.ptr, n * sizeof(uint32_t), path);
        restore(in, sp.absorbed.ptr, 7 * sizeof(unsigned long long), path);
        const bool has_ledger = read_pod<bool>(in, path);
        if (has_ledger) {
            if (!sp.ledger.ptr) sp.ledger.alloc(3);
            restore(in, sp.ledger.ptr, 3 * sizeof(double), path);
        } else if (sp.ledger.ptr) {
            AUR_HIP(hipMemset(sp.ledger.ptr, 0, 3 * sizeof(double)));
        }
        sp.ledger_pending = false;
        sp.slots = slots;
        sp.next_uid = next_uid;
        sp.tiled = false;           // the next sort rebuilds the tile index
        sp.pushes_since_sort = 0;
        sp.acc_next_clear = false;
    }
    const int models = read_pod<int>(in, path);
    AUR_REQUIRE(models == c.mcc_models(), AURORA_ERR_INVALID_ARGUMENT,
                "checkpoint collision models do not match this context");
    for (int model = 0; model < models; ++model) {
        restore(in, c.mcc_model(model).counts.ptr, (kMaxChannels + 3) * sizeof(unsigned long long),
                path);
        restore(in, c.mcc_model(model).energy_change.ptr, kMaxChannels * sizeof(double), path);
    }
    c.time = time;
    c.step_count = step;
    // the fields of the saved step: deposit the restored particles and solve
    // at the saved time (the electrode potentials of that time in 1D)
    deposit_all(c);
    convert_rho(c);
    solve_fields(c, electrode(c.left4, c.time), electrode(c.right4, c.time), true, true);
    AUR_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace aurora

using namespace aurora;

extern "C" {

int aurora_hip_checkpoint_save(aurora_hip_ctx* h, const char* path) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(path && *path, AURORA_ERR_INVALID_ARGUMENT, "checkpoint path is empty");
        checkpoint_save(c, path);
    });
}

int aurora_hip_checkpoint_load(aurora_hip_ctx* h, const char* path) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(path && *path, AURORA_ERR_INVALID_ARGUMENT, "checkpoint path is empty");
        checkpoint_load(c, path);
    });
}

int aurora_hip_physics_signature(aurora_hip_ctx* h, uint64_t* signature) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(signature, AURORA_ERR_INVALID_ARGUMENT, "null signature pointer");
        *signature = physics_signature(c);
    });
}

}  // extern "C"
