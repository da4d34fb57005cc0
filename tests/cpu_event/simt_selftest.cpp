// Checks of the 64-lane stand-in (include_simt/hip/hip_runtime.h) and of the
// event path's wavefront-level pieces under it:
//   ok:                  shuffle reduction, ballots of a divergent subset, the
//                        bank append of event.hpp (BankSink::push) with mixed
//                        destinations from divergent lanes, and the refill
//                        reservation (Refill::take) shared by the idle lanes
//                        of every wavefront: every particle lands once, every
//                        bank slot is handed out once;
//   divergent_uniform:   the two halves of a wavefront at whole-wavefront
//                        votes of different branches must stop the run;
//   readlane_inactive:   a readlane of a lane outside the executing group
//                        must stop the run.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../gegs_amd/csrc/event.hpp"

namespace gegs {
float lds_env[1];
}
using namespace gegs;

static int failures = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            failures++;                                             \
        }                                                           \
    } while (0)

static uint32_t sums[256];
static uint64_t subset_ballot[256];

void k_reduce(uint32_t *out, uint64_t *ballots) {
    const uint32_t t = threadIdx.x;
    out[t] = wave_sum(t + 1);
    // lanes 3 mod 4 take a branch and vote there, as a subset
    if ((t & 3u) == 3u) {
        GEGS_WAVE_SUBSET_OP(9);
        ballots[t] = __ballot(1);
    }
}

void k_divergent_uniform(int *out) {
    // whole-wavefront votes from the two sides of a branch
    if (threadIdx.x & 1u) out[threadIdx.x] = __any(1);
    else out[threadIdx.x] = (int)__popcll(__ballot(1));
}

void k_readlane_inactive(uint32_t *out) {
    if (threadIdx.x < 8) {
        GEGS_WAVE_SUBSET_OP(9);
        out[threadIdx.x] = __builtin_amdgcn_readlane(threadIdx.x, 40);   // lane 40 is not in the group
    }
}

// every lane with an odd thread index pushes one particle whose event
// depends on the lane, from divergent code
void k_push(EvArgs a) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    BankSink sink{&a, 0, 0u, true, true};
    if (gid & 1u) {
        Part p{};
        p.e = (float)gid;
        p.status = (int)(gid % 7u) == 0 ? P_PHOTON_STEP : (int)(gid % 3u) + P_COMPTON;
        p.job = 0;
        sink.hist = gid;
        sink.push(p);
    }
}

// lanes take bank slots until the bank is exhausted, each lane keeping a
// taken slot for (gid % 5) + 1 iterations
void k_refill(uint32_t *taken, uint32_t n, uint32_t *hits, uint32_t min_idle) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    Refill rf;
    rf.init();
    bool have = false;
    uint32_t left = 0;
    for (;;) {
        const uint32_t slot = rf.take(taken, n, !have, min_idle, 64);
        if (!have && slot != 0xFFFFFFFFu) {
            if (slot >= n) {
                fprintf(stderr, "slot %u beyond %u\n", slot, n);
                abort();
            }
            hits[slot]++;
            have = true;
            left = gid % 5u + 1u;
        }
        if (!__any(have)) break;
        if (have && --left == 0) have = false;
    }
}

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "ok";
    if (mode == "divergent_uniform") {
        int out[64] = {};
        cpu_launch(k_divergent_uniform, dim3(1), dim3(64), out);
        return 0;   // not reached: the stand-in stops the run
    }
    if (mode == "readlane_inactive") {
        uint32_t out[64] = {};
        cpu_launch(k_readlane_inactive, dim3(1), dim3(64), out);
        return 0;
    }
    // shuffle reduction and subset ballots
    cpu_launch(k_reduce, dim3(1), dim3(256), sums, subset_ballot);
    for (uint32_t t = 0; t < 256; ++t) {
        const uint32_t w = t / 64;
        CHECK(sums[t] == 64u * (64u * w) + 64u * 65u / 2u);
        if ((t & 3u) == 3u) CHECK(subset_ballot[t] == 0x8888888888888888ull);
    }
    // grouped appends with mixed destinations
    {
        const uint32_t cap = 64 * QSEG;
        EvState *st = (EvState *)aligned_alloc(alignof(EvState), sizeof(EvState));
        memset(st, 0, sizeof(EvState));
        std::vector<float> bank((size_t)cap * N_SLOTS * EVW, -1.0f);
        EvArgs a{};
        a.bank = bank.data();
        a.cap = cap;
        a.seg_cap = cap / QSEG;
        a.st = st;
        const uint32_t blocks = 5, threads = 192;
        cpu_launch(k_push, dim3(blocks), dim3(threads), a);
        std::vector<int> seen(blocks * threads, 0);
        uint32_t pushed = 0;
        for (int e = 0; e < N_SLOTS; ++e)
            for (int g = 0; g < QSEG; ++g) {
                const uint32_t c = st->count[e][g].v;
                CHECK(c <= a.seg_cap);
                for (uint32_t i = 0; i < c; ++i) {
                    const uint32_t idx = g * a.seg_cap + i;
                    const uint32_t gid = (uint32_t)bank[((size_t)e * EVW + 0) * cap + idx];
                    CHECK(gid < blocks * threads && (gid & 1u));
                    const int want = (int)(gid % 7u) == 0 ? P_PHOTON_STEP : (int)(gid % 3u) + P_COMPTON;
                    CHECK(e == want);
                    CHECK((uint32_t)g == (gid / threads) % QSEG);
                    seen[gid]++;
                    pushed++;
                }
            }
        for (uint32_t gid = 0; gid < blocks * threads; ++gid) CHECK(seen[gid] == (int)(gid & 1u));
        CHECK(pushed == blocks * threads / 2);
        CHECK(st->overflow == 0);
        free(st);
    }
    // refill reservations: every slot exactly once, for two thresholds
    for (uint32_t min_idle : {1u, 16u}) {
        const uint32_t n = 10007;
        std::vector<uint32_t> hits(n, 0);
        uint32_t taken = 0;
        cpu_launch(k_refill, dim3(3), dim3(256), &taken, n, hits.data(), min_idle);
        for (uint32_t i = 0; i < n; ++i) CHECK(hits[i] == 1u);
    }
    printf("simt self-test: %d failures, %llu wavefront operations\n", failures, (unsigned long long)simt::S.ops);
    return failures ? 1 : 0;
}
