// mcc.hip - null-collision MCC (Collision.cpp:831-993) and the 1D
// collision step with ionization products (Simulation.cpp:771-1006).
#include "aurora_internal.hpp"

namespace aurora {

void mcc_step(Context& c, double dt) {
    (void)dt;
    throw Failure{AURORA_ERR_UNSUPPORTED, "MCC step not built yet"};
}

}  // namespace aurora

using namespace aurora;

extern "C" {

int aurora_hip_mcc_load(aurora_hip_ctx* h, int, int, const int*, const double*, const int64_t*,
                        const double* const*, const double* const*, double, double, int64_t,
                        double, double, const int*, uint64_t) {
    return guarded(h, [](Context&) {
        throw Failure{AURORA_ERR_UNSUPPORTED, "MCC not built yet"};
    });
}

int aurora_hip_mcc_stats(aurora_hip_ctx* h, uint64_t*, double*) {
    return guarded(h, [](Context&) {
        throw Failure{AURORA_ERR_UNSUPPORTED, "MCC not built yet"};
    });
}

int aurora_hip_mcc_collide(aurora_hip_ctx* h, int64_t, double*, double, uint64_t, uint32_t,
                           uint64_t*, double*) {
    return guarded(h, [](Context&) {
        throw Failure{AURORA_ERR_UNSUPPORTED, "MCC not built yet"};
    });
}

}  // extern "C"
