// swarm.hip - batched homogeneous electron swarms (Swarm.cpp:940-1250).
#include "aurora_internal.hpp"

namespace aurora {
void swarm_destroy(Context& c) { c.swarm = nullptr; }
}  // namespace aurora

using namespace aurora;

extern "C" {

int aurora_hip_swarm_gas(aurora_hip_ctx* h, int, const int*, const double*, const int64_t*,
                         const double* const*, const double* const*, double, double, int64_t,
                         double) {
    return guarded(h, [](Context&) { throw Failure{AURORA_ERR_UNSUPPORTED, "swarm not built yet"}; });
}
int aurora_hip_swarm_submit(aurora_hip_ctx* h, const aurora_hip_swarm_request*, uint64_t, int64_t*) {
    return guarded(h, [](Context&) { throw Failure{AURORA_ERR_UNSUPPORTED, "swarm not built yet"}; });
}
int aurora_hip_swarm_advance(aurora_hip_ctx* h, int64_t, int64_t*) {
    return guarded(h, [](Context&) { throw Failure{AURORA_ERR_UNSUPPORTED, "swarm not built yet"}; });
}
int aurora_hip_swarm_poll(aurora_hip_ctx* h, int64_t, int*, aurora_hip_swarm_result*) {
    return guarded(h, [](Context&) { throw Failure{AURORA_ERR_UNSUPPORTED, "swarm not built yet"}; });
}
uint64_t aurora_hip_swarm_footprint(const aurora_hip_swarm_request* r) {
    return r ? static_cast<uint64_t>(r->particles) * 12u : 0u;
}

}  // extern "C"
