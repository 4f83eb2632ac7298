// api_core.hip - C ABI entry points: context, mesh, species state, fields,
// step orchestration.  Kernels live in particles.hip, poisson.hip, mcc.hip,
// swarm.hip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <new>

#include "aurora_internal.hpp"

using namespace aurora;

namespace aurora {

void flush_timers(Context& ctx) {
    if (ctx.pending.empty()) return;
    AUR_HIP(hipStreamSynchronize(ctx.stream));
    for (auto& p : ctx.pending) {
        float ms = 0.f;
        AUR_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        auto& t = ctx.timers[p.first];
        t.total_ms += ms;
        t.launches += 1;
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    ctx.pending.clear();
}

void check_device_error(Context& ctx) {
    uint32_t err = 0;
    AUR_HIP(hipMemcpyAsync(&err, ctx.dev_error.ptr, sizeof err, hipMemcpyDeviceToHost,
                           ctx.stream));
    AUR_HIP(hipStreamSynchronize(ctx.stream));
    if (!err) return;
    uint32_t zero = 0;
    AUR_HIP(hipMemcpy(ctx.dev_error.ptr, &zero, sizeof zero, hipMemcpyHostToDevice));
    if (err & DEV_ERR_B_RANGE) {
        const char axis = "xyz"[ctx.baxis];
        throw Failure{AURORA_ERR_B_OUT_OF_RANGE,
                      std::string("tabulated vector field query on axis ") + axis +
                          " lies outside the table range"};
    }
    if (err & DEV_ERR_RATE)
        throw Failure{AURORA_ERR_RATE_ABOVE_BOUND,
                      "MCC total collision frequency exceeds configured max_frequency"};
    if (err & DEV_ERR_CANDIDATES)
        throw Failure{AURORA_ERR_CANDIDATE_LIMIT,
                      "MCC candidate limit exceeded; reduce dt or increase "
                      "max_candidates_per_particle"};
    if (err & DEV_ERR_CAPACITY)
        throw Failure{AURORA_ERR_CAPACITY,
                      "reactive collisions exceeded max_particles_per_species"};
    if (err & DEV_ERR_NONFINITE)
        throw Failure{AURORA_ERR_INVALID_ARGUMENT, "MCC particle velocity must be finite"};
    throw Failure{AURORA_ERR_HIP, "device reported error bits " + std::to_string(err)};
}

// A tile-sorted species is re-sorted only when enough of its pushes since
// the last sort fell outside their tile's LDS box (those take the global
// gather/atomic path); results do not depend on the order either way.
bool sort_needed(Context& c, Species& sp) {
    if (!sp.tiled || c.sort_threshold <= 0.0 || sp.pushes_since_sort == 0 || sp.slots == 0)
        return true;
    unsigned long long out = 0;
    AUR_HIP(hipMemcpyAsync(&out, sp.absorbed.ptr + 6, sizeof out, hipMemcpyDeviceToHost,
                           c.stream));
    AUR_HIP(hipStreamSynchronize(c.stream));
    const double frac = static_cast<double>(out) /
                        (static_cast<double>(sp.pushes_since_sort) * static_cast<double>(sp.slots));
    if (frac >= c.sort_threshold) return true;
    return false;
}

// The renewal rule: with a cycle cost C(n) = sort + pushes 1..n and pushes
// that slow down as the cycle ages, the long-run cost per step C(n)/n is
// least at the first n whose push exceeds C(n)/n.  The push compared is the
// mean of the last three timed pushes (timing noise).  Timings are read one
// step late (the previous step's push has completed by then), so the
// decision lags by one step and the host never stalls the queue.
bool adaptive_sort_due(Context& c, Species& sp) {
    if (sp.precision != 32 || !c.tiled_push_ok()) return false;  // no tiled push to speed up
    if (!sp.tiled) return sp.slots > 0;
    if (sp.sort_ev.recorded && hipEventQuery(sp.sort_ev.b) == hipSuccess)
        sp.cycle_ms += sp.sort_ev.take_ms();
    EventPair& ev = sp.push_ev[sp.push_ev_cur];   // recorded by the previous step
    if (ev.recorded) {
        const double p = ev.take_ms();
        sp.cycle_ms += p;
        sp.recent_ms[2] = sp.recent_ms[1];
        sp.recent_ms[1] = sp.recent_ms[0];
        sp.recent_ms[0] = p;
        ++sp.timed_pushes;
    }
    if (c.sort_max_interval > 0 && sp.pushes_since_sort >= c.sort_max_interval) return true;
    if (sp.sort_ev.recorded || sp.timed_pushes < 3) return false;
    const double recent = (sp.recent_ms[0] + sp.recent_ms[1] + sp.recent_ms[2]) / 3.0;
    return recent * static_cast<double>(sp.timed_pushes) > sp.cycle_ms;
}

double electrode(const double* p, double t) {
    // Simulation.cpp:678-685
    return p[0] + p[1] * std::sin(2.0 * 3.14159265358979323846 * p[2] * t + p[3]);
}

bool any_fp64(const Context& c) {
    for (const auto& sp : c.species)
        if (sp.precision == 64) return true;
    return false;
}

void ensure_fields(Context& c, bool phi, bool e64) {
    if ((phi && !c.phi_valid) || (e64 && !c.e64_valid)) {
        ensure_rho(c);
        solve_fields(c, electrode(c.left4, c.time), electrode(c.right4, c.time), true, true);
    }
}

}  // namespace aurora

namespace {

bool finite_positive(double v) { return std::isfinite(v) && v > 0.0; }

// Locate a coordinate on one axis the way the reference's deposit and
// gather do (wrap_periodic or clamp, floor, i = min(i, n-2) on Dirichlet
// axes), then round the offset to the storage precision keeping it in
// [0,1) (periodic) or [0,1] (Dirichlet).
template <typename R>
void locate_host(double x, int64_t n, double L, double d, int bc, int32_t& cell, R& off) {
    if (bc == AURORA_PERIODIC) {
        const double xp = std::fmod(std::fmod(x, L) + L, L);
        const double g = xp / d;
        int64_t i = static_cast<int64_t>(std::floor(g));
        double f = g - static_cast<double>(i);
        R fr = static_cast<R>(f);
        if (fr >= R(1)) { fr = R(0); i += 1; }
        i %= n;
        cell = static_cast<int32_t>(i);
        off = fr;
    } else {
        const double xp = std::min(std::max(x, 0.0), L);
        const double g = xp / d;
        double fl = std::floor(g);
        fl = std::min(fl, static_cast<double>(n - 2));
        const int64_t i = static_cast<int64_t>(fl);
        double f = g - static_cast<double>(i);
        f = std::min(std::max(f, 0.0), 1.0);
        cell = static_cast<int32_t>(i);
        off = static_cast<R>(f);
    }
}

int64_t unpack_axis(int32_t c, int dim, int a) {
    if (dim == 1) return c;
    const int shift = pack_shift(dim);
    return (c >> (shift * a)) & ((1 << shift) - 1);
}

template <typename R>
void upload_species(Context& ctx, Species& sp, int64_t count, const double* pos,
                    const double* vel, const uint8_t* alive) {
    const int dim = ctx.dim;
    std::vector<int32_t> cell(static_cast<size_t>(count));
    std::vector<R> off[3], v[3];
    for (int a = 0; a < 3; ++a) {
        off[a].assign(static_cast<size_t>(count), R(0));
        v[a].assign(static_cast<size_t>(count), R(0));
    }
    std::vector<uint32_t> uid(static_cast<size_t>(count));
    for (int64_t p = 0; p < count; ++p) {
        uid[p] = static_cast<uint32_t>(p);
        for (int a = 0; a < 3; ++a) v[a][p] = static_cast<R>(vel[3 * p + a]);
        if (alive && !alive[p]) {
            cell[p] = -1;
            continue;
        }
        int64_t packed = 0;
        const int shift = pack_shift(dim);
        for (int a = 0; a < dim; ++a) {
            int32_t c;
            R o;
            locate_host<R>(pos[dim * p + a], ctx.n[a], ctx.L[a], ctx.d[a], ctx.bc, c, o);
            off[a][p] = o;
            packed |= static_cast<int64_t>(c) << (shift * a);
        }
        cell[p] = static_cast<int32_t>(packed);
    }
    AUR_HIP(hipMemcpy(sp.cell.ptr, cell.data(), count * sizeof(int32_t), hipMemcpyHostToDevice));
    AUR_HIP(hipMemcpy(sp.uid.ptr, uid.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (int a = 0; a < 3; ++a) {
        AUR_HIP(hipMemcpy(sp.vel[a].ptr, v[a].data(), count * sizeof(R), hipMemcpyHostToDevice));
        if (a < dim)
            AUR_HIP(hipMemcpy(sp.off[a].ptr, off[a].data(), count * sizeof(R),
                              hipMemcpyHostToDevice));
    }
}

template <typename R>
void download_species(Context& ctx, Species& sp, double* pos, double* vel, uint8_t* alive,
                      uint32_t* uid) {
    const int64_t count = sp.slots;
    const int dim = ctx.dim;
    std::vector<int32_t> cell(static_cast<size_t>(count));
    AUR_HIP(hipStreamSynchronize(ctx.stream));
    AUR_HIP(hipMemcpy(cell.data(), sp.cell.ptr, count * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (uid)
        AUR_HIP(hipMemcpy(uid, sp.uid.ptr, count * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<R> buf(static_cast<size_t>(count));
    if (vel) {
        for (int a = 0; a < 3; ++a) {
            AUR_HIP(hipMemcpy(buf.data(), sp.vel[a].ptr, count * sizeof(R),
                              hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < count; ++p) vel[3 * p + a] = static_cast<double>(buf[p]);
        }
    }
    if (pos) {
        for (int a = 0; a < dim; ++a) {
            AUR_HIP(hipMemcpy(buf.data(), sp.off[a].ptr, count * sizeof(R),
                              hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < count; ++p) {
                if (cell[p] < 0) { pos[dim * p + a] = 0.0; continue; }
                const int64_t idx = unpack_axis(cell[p], dim, a);
                pos[dim * p + a] =
                    (static_cast<double>(idx) + static_cast<double>(buf[p])) * ctx.d[a];
            }
        }
    }
    if (alive)
        for (int64_t p = 0; p < count; ++p) alive[p] = cell[p] >= 0 ? 1 : 0;
}

}  // namespace

extern "C" {

int aurora_hip_create(int device, aurora_hip_ctx** out) {
    if (!out) return AURORA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto* h = new (std::nothrow) aurora_hip_ctx();
    if (!h) return AURORA_ERR_HIP;
    int rc = guarded(h, [&](Context& c) {
        int ndev = 0;
        AUR_HIP(hipGetDeviceCount(&ndev));
        AUR_REQUIRE(device >= 0 && device < ndev, AURORA_ERR_INVALID_ARGUMENT,
                    "HIP device ordinal out of range");
        c.device = device;
        AUR_HIP(hipSetDevice(device));
        AUR_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        AUR_HIP(hipGetDeviceProperties(&prop, device));
        c.cu_count = prop.multiProcessorCount;
        char buf[256];
        std::snprintf(buf, sizeof buf, "%s, %d CUs, %.1f GiB HBM", prop.gcnArchName,
                      prop.multiProcessorCount,
                      static_cast<double>(prop.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0));
        c.info = buf;
        c.dev_error.alloc(1);
        AUR_HIP(hipMemset(c.dev_error.ptr, 0, sizeof(uint32_t)));
        collide_scratch_init(device);
    });
    if (rc != AURORA_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return AURORA_OK;
}

int aurora_hip_destroy(aurora_hip_ctx* h) {
    if (!h) return AURORA_OK;
    (void)hipSetDevice(h->c.device);
    if (h->c.stream) (void)hipStreamSynchronize(h->c.stream);
    swarm_destroy(h->c);
    ensembles_destroy(h->c);
    for (auto& p : h->c.pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    h->c.pending.clear();
    hipStream_t s = h->c.stream;
    delete h;  // releases device buffers
    if (s) (void)hipStreamDestroy(s);
    return AURORA_OK;
}

const char* aurora_hip_last_error(const aurora_hip_ctx* h) {
    return h ? h->c.last_error.c_str() : "null context";
}

int aurora_hip_synchronize(aurora_hip_ctx* h) {
    return guarded(h, [](Context& c) {
        AUR_HIP(hipStreamSynchronize(c.stream));
        check_device_error(c);
    });
}

int aurora_hip_device_info(aurora_hip_ctx* h, char* buf, size_t len) {
    return guarded(h, [&](Context& c) {
        if (buf && len) std::snprintf(buf, len, "%s", c.info.c_str());
    });
}

int aurora_hip_mesh_bind(aurora_hip_ctx* h, int dim, const int64_t* n, const double* length,
                         int boundary, double permittivity, double depth) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(dim >= 1 && dim <= 3 && n && length, AURORA_ERR_INVALID_ARGUMENT,
                    "mesh dimension must be 1, 2 or 3");
        AUR_REQUIRE(boundary == AURORA_PERIODIC || boundary == AURORA_DIRICHLET,
                    AURORA_ERR_INVALID_ARGUMENT, "unknown field boundary");
        AUR_REQUIRE(finite_positive(permittivity), AURORA_ERR_INVALID_ARGUMENT,
                    "field-solver permittivity must be positive and finite");
        if (dim == 2)
            AUR_REQUIRE(finite_positive(depth), AURORA_ERR_INVALID_ARGUMENT,
                        "2D deposition depth must be positive and finite");
        int64_t nodes = 1;
        for (int a = 0; a < dim; ++a) {
            AUR_REQUIRE(n[a] >= 3, AURORA_ERR_INVALID_ARGUMENT, "mesh extents must be at least 3");
            AUR_REQUIRE(finite_positive(length[a]), AURORA_ERR_INVALID_ARGUMENT,
                        "domain length must be positive and finite");
            nodes *= n[a];
        }
        AUR_REQUIRE(nodes < (int64_t{1} << 31), AURORA_ERR_UNSUPPORTED,
                    "mesh node count exceeds the int32 cell index of this build");
        for (int a = 0; a < dim; ++a)
            AUR_REQUIRE(dim == 1 || n[a] <= (int64_t{1} << pack_shift(dim)), AURORA_ERR_UNSUPPORTED,
                        "mesh extent exceeds the packed cell index (3D: 1024, 2D: 32768 per axis)");
        AUR_HIP(hipSetDevice(c.device));
        AUR_HIP(hipStreamSynchronize(c.stream));
        c.species.clear();
        c.mcc.loaded = false;
        c.mcc_more.clear();
        c.dim = dim;
        for (int a = 0; a < 3; ++a) {
            c.n[a] = a < dim ? n[a] : 1;
            c.L[a] = a < dim ? length[a] : 1.0;
            c.d[a] = a < dim ? length[a] / static_cast<double>(
                                               boundary == AURORA_PERIODIC ? n[a] : n[a] - 1)
                             : 1.0;
        }
        c.bc = boundary;
        c.eps = permittivity;
        c.depth = dim == 2 ? depth : 1.0;
        c.nnodes = nodes;
        for (int f = 0; f < 6; ++f)
            c.pb[f] = boundary == AURORA_PERIODIC ? AURORA_PB_PERIODIC : AURORA_PB_ABSORB;
        c.rho.alloc(nodes);
        c.phi.alloc(nodes);
        c.E64.alloc(static_cast<size_t>(nodes) * dim);
        c.E32.alloc(static_cast<size_t>(nodes) * (dim == 3 ? 4 : dim));
        AUR_HIP(hipMemset(c.rho.ptr, 0, nodes * sizeof(double)));
        AUR_HIP(hipMemset(c.phi.ptr, 0, nodes * sizeof(double)));
        AUR_HIP(hipMemset(c.E64.ptr, 0, nodes * dim * sizeof(double)));
        AUR_HIP(hipMemset(c.E32.ptr, 0, c.E32.count * sizeof(float)));
        c.time = 0.0;
        c.step_count = 0;
        c.phi_valid = false;
        prepare_solver(c);
        if (c.has_profile) set_b_profile_device(c);
    });
}

int aurora_hip_particle_boundaries(aurora_hip_ctx* h, const int* policy) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(policy && c.dim > 0, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
        for (int f = 0; f < 2 * c.dim; ++f) {
            AUR_REQUIRE(policy[f] >= AURORA_PB_ABSORB && policy[f] <= AURORA_PB_PERIODIC,
                        AURORA_ERR_INVALID_ARGUMENT, "unresolved particle boundary policy");
            AUR_REQUIRE(!(policy[f] == AURORA_PB_PERIODIC) || c.bc == AURORA_PERIODIC ||
                            c.dim == 1,
                        AURORA_ERR_UNSUPPORTED,
                        "periodic particles on a Dirichlet mesh are not part of this build");
            c.pb[f] = policy[f];
        }
    });
}

int aurora_hip_species_add(aurora_hip_ctx* h, double charge, double mass, double weight,
                           int64_t capacity, int precision, int* species_id) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
        AUR_REQUIRE(std::isfinite(charge) && finite_positive(mass) && finite_positive(weight),
                    AURORA_ERR_INVALID_ARGUMENT,
                    "species charge must be finite, mass and weight positive and finite");
        AUR_REQUIRE(capacity > 0 && capacity < (int64_t{1} << 31), AURORA_ERR_INVALID_ARGUMENT,
                    "species capacity must be positive and below 2^31");
        AUR_REQUIRE(precision == 32 || precision == 64, AURORA_ERR_INVALID_ARGUMENT,
                    "species precision must be 32 or 64");
        AUR_HIP(hipSetDevice(c.device));
        c.species.emplace_back();
        Species& sp = c.species.back();
        sp.charge = charge;
        sp.mass = mass;
        sp.weight = weight;
        sp.precision = precision;
        sp.capacity = capacity;
        const size_t e = sp.elem();
        sp.cell.alloc(capacity);
        sp.uid.alloc(capacity);
        for (int a = 0; a < 3; ++a) {
            sp.vel[a].alloc(capacity * e);
            if (a < c.dim) sp.off[a].alloc(capacity * e);
        }
        sp.acc.alloc(c.nnodes);
        AUR_HIP(hipMemset(sp.acc.ptr, 0, c.nnodes * sizeof(int64_t)));
        // 6 face counters + pushes that left their tile's LDS box
        sp.absorbed.alloc(7);
        AUR_HIP(hipMemset(sp.absorbed.ptr, 0, 7 * sizeof(unsigned long long)));
        if (species_id) *species_id = static_cast<int>(c.species.size() - 1);
    });
}

int aurora_hip_species_upload(aurora_hip_ctx* h, int sid, int64_t count, const double* position,
                              const double* velocity, const uint8_t* alive, int velocity_kind) {
    return guarded(h, [&](Context& c) {
        Species& sp = species_at(c, sid);
        AUR_REQUIRE(count >= 0 && count <= sp.capacity, AURORA_ERR_CAPACITY,
                    "particle count exceeds max_particles_per_species");
        AUR_REQUIRE(count == 0 || (position && velocity), AURORA_ERR_INVALID_ARGUMENT,
                    "position and velocity arrays are required");
        (void)velocity_kind;
        AUR_HIP(hipSetDevice(c.device));
        AUR_HIP(hipStreamSynchronize(c.stream));
        for (int64_t p = 0; p < count * c.dim; ++p)
            AUR_REQUIRE(std::isfinite(position[p]), AURORA_ERR_INVALID_ARGUMENT,
                        "particle positions must be finite");
        if (count) {
            if (sp.precision == 64) upload_species<double>(c, sp, count, position, velocity, alive);
            else upload_species<float>(c, sp, count, position, velocity, alive);
        }
        sp.slots = count;
        sp.next_uid = static_cast<uint32_t>(count);
        sp.tiled = false;
    });
}

int aurora_hip_species_count(aurora_hip_ctx* h, int sid, int64_t* slots, int64_t* live) {
    return guarded(h, [&](Context& c) {
        Species& sp = species_at(c, sid);
        if (slots) *slots = sp.slots;
        if (live) *live = live_count(c, sp);
    });
}

int aurora_hip_species_download(aurora_hip_ctx* h, int sid, int64_t capacity, double* position,
                                double* velocity_half, uint8_t* alive, uint32_t* uid,
                                int64_t* slots) {
    return guarded(h, [&](Context& c) {
        Species& sp = species_at(c, sid);
        if (slots) *slots = sp.slots;
        AUR_REQUIRE(capacity >= sp.slots, AURORA_ERR_INVALID_ARGUMENT,
                    "output arrays are smaller than the species slot count");
        if (sp.slots == 0) return;
        if (sp.precision == 64) download_species<double>(c, sp, position, velocity_half, alive, uid);
        else download_species<float>(c, sp, position, velocity_half, alive, uid);
    });
}

int aurora_hip_species_download_raw(aurora_hip_ctx* h, int sid, int64_t capacity, int32_t* cell,
                                    double* offsets, double* velocity_half, int64_t* slots) {
    return guarded(h, [&](Context& c) {
        Species& sp = species_at(c, sid);
        if (slots) *slots = sp.slots;
        AUR_REQUIRE(capacity >= sp.slots, AURORA_ERR_INVALID_ARGUMENT,
                    "output arrays are smaller than the species slot count");
        const int64_t n = sp.slots;
        if (n == 0) return;
        AUR_HIP(hipStreamSynchronize(c.stream));
        if (cell) {
            AUR_HIP(hipMemcpy(cell, sp.cell.ptr, n * sizeof(int32_t), hipMemcpyDeviceToHost));
            // packed corner coordinates -> linear node index (k ny + j) nx + i
            for (int64_t p = 0; p < n; ++p) {
                if (cell[p] < 0) continue;
                int64_t lin = 0, stride = 1;
                for (int a = 0; a < c.dim; ++a) {
                    lin += unpack_axis(cell[p], c.dim, a) * stride;
                    stride *= c.n[a];
                }
                cell[p] = static_cast<int32_t>(lin);
            }
        }
        std::vector<uint8_t> buf(static_cast<size_t>(n) * sp.elem());
        auto widen = [&](const DevBuf<uint8_t>& src, double* dst, int stride, int col) {
            AUR_HIP(hipMemcpy(buf.data(), src.ptr, buf.size(), hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < n; ++p) {
                double v;
                if (sp.precision == 64) std::memcpy(&v, buf.data() + 8 * p, 8);
                else {
                    float f;
                    std::memcpy(&f, buf.data() + 4 * p, 4);
                    v = static_cast<double>(f);
                }
                dst[stride * p + col] = v;
            }
        };
        if (offsets)
            for (int a = 0; a < c.dim; ++a) widen(sp.off[a], offsets, c.dim, a);
        if (velocity_half)
            for (int a = 0; a < 3; ++a) widen(sp.vel[a], velocity_half, 3, a);
    });
}

int aurora_hip_species_charge_fixed(aurora_hip_ctx* h, int sid, int64_t* acc) {
    return guarded(h, [&](Context& c) {
        Species& sp = species_at(c, sid);
        AUR_REQUIRE(acc, AURORA_ERR_INVALID_ARGUMENT, "null output");
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(acc, sp.acc.ptr, c.nnodes * sizeof(int64_t), hipMemcpyDeviceToHost));
    });
}

int aurora_hip_set_uniform_b(aurora_hip_ctx* h, const double* b3) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(b3 && std::isfinite(b3[0]) && std::isfinite(b3[1]) && std::isfinite(b3[2]),
                    AURORA_ERR_INVALID_ARGUMENT, "magnetic_field components must be finite");
        AUR_REQUIRE(!(c.has_profile && (b3[0] != 0 || b3[1] != 0 || b3[2] != 0)),
                    AURORA_ERR_INVALID_ARGUMENT,
                    "uniform magnetic_field components and magnetic_field_profile are "
                    "mutually exclusive");
        for (int a = 0; a < 3; ++a) c.Buni[a] = b3[a];
    });
}

int aurora_hip_set_b_profile(aurora_hip_ctx* h, int axis, int64_t rows, const double* coords,
                             const double* values) {
    return guarded(h, [&](Context& c) {
        if (rows == 0) {
            c.has_profile = false;
            c.brows = 0;
            return;
        }
        AUR_REQUIRE(axis >= 0 && axis <= 2, AURORA_ERR_INVALID_ARGUMENT,
                    "coordinate axis must be x, y, or z");
        AUR_REQUIRE(rows >= 2 && coords && values, AURORA_ERR_INVALID_ARGUMENT,
                    "tabulated vector field requires at least two rows");
        for (int64_t i = 0; i < rows; ++i) {
            AUR_REQUIRE(std::isfinite(coords[i]) && std::isfinite(values[3 * i]) &&
                            std::isfinite(values[3 * i + 1]) && std::isfinite(values[3 * i + 2]),
                        AURORA_ERR_INVALID_ARGUMENT, "tabulated vector field values must be finite");
            AUR_REQUIRE(i == 0 || coords[i] > coords[i - 1], AURORA_ERR_INVALID_ARGUMENT,
                        "tabulated vector field coordinates must be strictly increasing");
        }
        AUR_REQUIRE(c.Buni[0] == 0 && c.Buni[1] == 0 && c.Buni[2] == 0,
                    AURORA_ERR_INVALID_ARGUMENT,
                    "uniform magnetic_field components and magnetic_field_profile are "
                    "mutually exclusive");
        if (c.dim > 0 && axis < c.dim) {
            // validate_domain (PrescribedField.cpp:115-134)
            AUR_REQUIRE(coords[0] <= 0.0 && coords[rows - 1] >= c.L[axis],
                        AURORA_ERR_INVALID_ARGUMENT,
                        "tabulated field does not cover the simulation domain");
        }
        c.has_profile = true;
        c.baxis = axis;
        c.brows = rows;
        c.bcoords.alloc(rows);
        c.bvalues.alloc(rows * 3);
        AUR_HIP(hipMemcpy(c.bcoords.ptr, coords, rows * sizeof(double), hipMemcpyHostToDevice));
        AUR_HIP(hipMemcpy(c.bvalues.ptr, values, rows * 3 * sizeof(double), hipMemcpyHostToDevice));
        if (c.dim > 0) set_b_profile_device(c);
    });
}

int aurora_hip_set_electrodes(aurora_hip_ctx* h, const double* left4, const double* right4) {
    return guarded(h, [&](Context& c) {
        for (int k = 0; k < 4; ++k) {
            c.left4[k] = left4 ? left4[k] : 0.0;
            c.right4[k] = right4 ? right4[k] : 0.0;
        }
    });
}

int aurora_hip_set_electric(aurora_hip_ctx* h, const double* E) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0 && E, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(c.E64.ptr, E, c.nnodes * c.dim * sizeof(double), hipMemcpyHostToDevice));
        c.e64_valid = true;
        std::vector<float> packed(c.E32.count, 0.f);
        const int stride = c.dim == 3 ? 4 : c.dim;
        for (int64_t i = 0; i < c.nnodes; ++i)
            for (int a = 0; a < c.dim; ++a)
                packed[i * stride + a] = static_cast<float>(E[a * c.nnodes + i]);
        AUR_HIP(hipMemcpy(c.E32.ptr, packed.data(), packed.size() * sizeof(float),
                          hipMemcpyHostToDevice));
    });
}

int aurora_hip_get_fields(aurora_hip_ctx* h, double* rho, double* phi, double* E) {
    return guarded(h, [&](Context& c) {
        if (rho || phi || E) ensure_rho(c);
        if (phi || E) ensure_fields(c, phi != nullptr, E != nullptr);
        AUR_HIP(hipStreamSynchronize(c.stream));
        if (rho) AUR_HIP(hipMemcpy(rho, c.rho.ptr, c.nnodes * sizeof(double), hipMemcpyDeviceToHost));
        if (phi) AUR_HIP(hipMemcpy(phi, c.phi.ptr, c.nnodes * sizeof(double), hipMemcpyDeviceToHost));
        if (E)
            AUR_HIP(hipMemcpy(E, c.E64.ptr, c.nnodes * c.dim * sizeof(double),
                              hipMemcpyDeviceToHost));
    });
}

int aurora_hip_set_rho(aurora_hip_ctx* h, const double* rho) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0 && rho, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(c.rho.ptr, rho, c.nnodes * sizeof(double), hipMemcpyHostToDevice));
        c.rho_valid = true;
    });
}

int aurora_hip_deposit(aurora_hip_ctx* h) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
        deposit_all(c);
        convert_rho(c);
        flush_timers(c);
    });
}

int aurora_hip_field_solve(aurora_hip_ctx* h, double phi_left, double phi_right, int want_phi) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0, AURORA_ERR_INVALID_ARGUMENT, "bind a mesh first");
        ensure_rho(c);
        solve_fields(c, phi_left, phi_right, want_phi != 0);
        flush_timers(c);
    });
}

int aurora_hip_push(aurora_hip_ctx* h, double dt) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0 && finite_positive(dt), AURORA_ERR_INVALID_ARGUMENT,
                    "simulation dt must be positive and finite");
        push_all(c, dt, false);
        check_device_error(c);
        flush_timers(c);
    });
}

int aurora_hip_initialize(aurora_hip_ctx* h, double dt) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0 && finite_positive(dt), AURORA_ERR_INVALID_ARGUMENT,
                    "simulation dt must be positive and finite");
        c.time = 0.0;
        c.step_count = 0;
        // the wall counters and 1D ledgers start over (Simulation.cpp:698-709)
        for (auto& sp : c.species) {
            AUR_HIP(hipMemsetAsync(sp.absorbed.ptr, 0, 6 * sizeof(unsigned long long), c.stream));
            if (sp.ledger.ptr)
                AUR_HIP(hipMemsetAsync(sp.ledger.ptr, 0, 3 * sizeof(double), c.stream));
            sp.ledger_pending = false;
        }
        deposit_all(c);
        convert_rho(c);
        solve_fields(c, electrode(c.left4, 0.0), electrode(c.right4, 0.0), true);
        init_half_all(c, dt);
        check_device_error(c);
        flush_timers(c);
    });
}

int aurora_hip_step(aurora_hip_ctx* h, double dt, int64_t nsteps) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(c.dim > 0 && finite_positive(dt), AURORA_ERR_INVALID_ARGUMENT,
                    "simulation dt must be positive and finite");
        AUR_REQUIRE(nsteps >= 0, AURORA_ERR_INVALID_ARGUMENT, "step count must be non-negative");
        // 1D MCC with products: the slot counts stay on the device for the
        // call (no host wait per step); the host's copies come back at its
        // end, before a sort, and on an error
        const bool dev_counts = nsteps > 0 && device_counts_apply(c);
        struct Leave {
            Context& c;
            ~Leave() {
                try {
                    device_counts_leave(c);
                } catch (...) {
                    c.counts_on_device = false;
                }
            }
        } leave{c};
        if (dev_counts) device_counts_enter(c);
        for (int64_t s = 0; s < nsteps; ++s) {
            // kick + drift + boundary + deposit at x^{n+1} (fused)
            push_all(c, dt, true, c.dim == 1);
            const double t1 = c.time + dt;
            solve_step_fields(c, electrode(c.left4, t1), electrode(c.right4, t1));
            if (c.dim == 1) ledger_work_1d(c, dt);
            if (c.mcc.loaded) mcc_step(c, dt);
            c.time = t1;
            c.step_count += 1;
            if (c.sort_policy == AURORA_SORT_ADAPTIVE) {
                for (auto& sp : c.species)
                    if (adaptive_sort_due(c, sp)) {
                        device_counts_leave(c);
                        sort_species(c, sp);
                    }
            } else if (c.sort_interval > 0 && c.step_count % c.sort_interval == 0) {
                device_counts_leave(c);
                for (auto& sp : c.species)
                    if (sort_needed(c, sp)) sort_species(c, sp);
            }
            if (dev_counts && !c.counts_on_device && s + 1 < nsteps) device_counts_enter(c);
        }
        device_counts_leave(c);
        check_device_error(c);
        flush_timers(c);
    });
}

int aurora_hip_time(aurora_hip_ctx* h, double* time, int64_t* step) {
    return guarded(h, [&](Context& c) {
        if (time) *time = c.time;
        if (step) *step = c.step_count;
    });
}

int aurora_hip_set_sort_threshold(aurora_hip_ctx* h, double fraction) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(std::isfinite(fraction) && fraction >= 0.0 && fraction <= 1.0,
                    AURORA_ERR_INVALID_ARGUMENT, "sort threshold must be in [0, 1]");
        c.sort_threshold = fraction;
    });
}

int aurora_hip_set_sort_policy(aurora_hip_ctx* h, int policy, int64_t max_interval) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(policy == AURORA_SORT_FIXED || policy == AURORA_SORT_ADAPTIVE,
                    AURORA_ERR_INVALID_ARGUMENT, "sort policy must be fixed or adaptive");
        AUR_REQUIRE(max_interval >= 0, AURORA_ERR_INVALID_ARGUMENT,
                    "sort max_interval must be non-negative");
        c.sort_policy = policy;
        c.sort_max_interval = max_interval;
    });
}

int aurora_hip_set_sort_method(aurora_hip_ctx* h, int method) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(method == AURORA_SORT_LOCAL || method == AURORA_SORT_RADIX,
                    AURORA_ERR_INVALID_ARGUMENT, "sort method must be local or radix");
        c.sort_method = method;
    });
}

int aurora_hip_set_sort_interval(aurora_hip_ctx* h, int64_t interval) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE(interval >= 0, AURORA_ERR_INVALID_ARGUMENT, "sort interval must be >= 0");
        c.sort_interval = interval;
    });
}

int aurora_hip_set_tiling(aurora_hip_ctx* h, int tx, int ty, int tz, int halo) {
    return guarded(h, [&](Context& c) {
        if (tx == 0) {
            c.tiled_enabled = false;
            return;
        }
        // 3D shapes, then 2D ones (z extent 1)
        static const int shapes[][4] = {{8, 8, 8, 2}, {8, 8, 8, 1}, {8, 8, 4, 2}, {8, 8, 4, 1}, {4, 4, 4, 2},
                                        {16, 8, 4, 2}, {8, 8, 1, 2}, {8, 8, 1, 1}, {16, 16, 1, 2},
                                        {16, 8, 1, 2}};
        bool known = false;
        for (const auto& s : shapes)
            known = known || (s[0] == tx && s[1] == ty && s[2] == tz && s[3] == halo);
        AUR_REQUIRE(known, AURORA_ERR_UNSUPPORTED, "tile shape is not built into this library");
        c.tiled_enabled = true;
        c.tile[0] = tx;
        c.tile[1] = ty;
        c.tile[2] = tz;
        c.halo = halo;
        for (auto& sp : c.species) sp.tiled = false;
    });
}

int aurora_hip_set_push_variant(aurora_hip_ctx* h, int block, int prefetch) {
    return guarded(h, [&](Context& c) {
        AUR_REQUIRE((block == 256 || block == 512) && prefetch >= 0 && prefetch <= 15,
                    AURORA_ERR_UNSUPPORTED, "push variant is not built into this library");
        c.push_block = block;
        c.push_prefetch = prefetch;
    });
}

int aurora_hip_sort(aurora_hip_ctx* h) {
    return guarded(h, [&](Context& c) {
        for (auto& sp : c.species) sort_species(c, sp);
        flush_timers(c);
    });
}

int aurora_hip_boundary_losses(aurora_hip_ctx* h, int sid, int64_t* absorbed6, double* energy2) {
    return guarded(h, [&](Context& c) {
        Species& sp = species_at(c, sid);
        unsigned long long counts[6];
        AUR_HIP(hipStreamSynchronize(c.stream));
        AUR_HIP(hipMemcpy(counts, sp.absorbed.ptr, sizeof counts, hipMemcpyDeviceToHost));
        if (absorbed6)
            for (int f = 0; f < 6; ++f) absorbed6[f] = static_cast<int64_t>(counts[f]);
        if (energy2) {
            energy2[0] = energy2[1] = 0.0;
            if (sp.ledger.ptr)
                AUR_HIP(hipMemcpy(energy2, sp.ledger.ptr, 2 * sizeof(double),
                                  hipMemcpyDeviceToHost));
        }
    });
}

int aurora_hip_power_transfer(aurora_hip_ctx* h, int sid, double* electric_work) {
    return guarded(h, [&](Context& c) {
        Species& sp = species_at(c, sid);
        AUR_REQUIRE(electric_work, AURORA_ERR_INVALID_ARGUMENT, "null electric_work pointer");
        *electric_work = 0.0;
        AUR_HIP(hipStreamSynchronize(c.stream));
        if (sp.ledger.ptr) {
            double led[3];
            AUR_HIP(hipMemcpy(led, sp.ledger.ptr, sizeof led, hipMemcpyDeviceToHost));
            *electric_work = led[2];
        }
    });
}

int aurora_hip_kernel_timing(aurora_hip_ctx* h, int enable) {
    return guarded(h, [&](Context& c) {
        flush_timers(c);
        c.timing = enable != 0;
        if (enable) c.timers.clear();
    });
}

int aurora_hip_kernel_time(aurora_hip_ctx* h, const char* name, double* total_ms,
                           int64_t* launches) {
    return guarded(h, [&](Context& c) {
        flush_timers(c);
        auto it = c.timers.find(name ? name : "");
        if (total_ms) *total_ms = it == c.timers.end() ? 0.0 : it->second.total_ms;
        if (launches) *launches = it == c.timers.end() ? 0 : it->second.launches;
    });
}

}  // extern "C"
