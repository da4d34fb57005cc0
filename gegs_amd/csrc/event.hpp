// event.hpp -- event-queued transport (placeholder until the queue kernels land)
#pragma once
#include "transport.hpp"

namespace gegs {

struct EventConfig {
    uint32_t block = 256;
    size_t capacity = 0;
};

struct EventBanks {
    hipError_t reserve(uint64_t, uint32_t, const EventConfig &) { return hipErrorNotSupported; }
    void release() {}
};

inline hipError_t run_event_batch(EventBanks &, const Env &, const gegs_source *, uint32_t, uint32_t, uint64_t,
                                  double *, double *, uint32_t *, unsigned long long *, const EventConfig &,
                                  hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace gegs
