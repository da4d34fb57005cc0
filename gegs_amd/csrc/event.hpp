// event.hpp -- event-queued transport: SoA particle banks in HBM, one kernel
// per event type, wavefront-aggregated queue appends.
//
// The reference already splits transport by event type (ten global AoS
// stacks, one kernel each, CUDA_EGS.h:362-367, launched with a host sync and
// a counter readback after every kernel, CUDA_EGS.cu:226-365).  Here:
//   * each event type has a bank of particles stored word-major
//     ([event][word][slot]), so the 64 lanes of a wavefront load and store
//     consecutive words;
//   * a kernel drains its bank and appends every particle it forwards or
//     creates to the bank of the next event type with one atomic per
//     wavefront and destination (ballot, popcount, lane rank), instead of one
//     contended atomic per particle (utility.cu:222-306);
//   * the photon free flight and the electron condensed-history step run as
//     persistent loops that refill idle lanes from the bank every iteration,
//     so lanes whose particle interacted or stopped take new work instead of
//     idling until the slowest lane of the wavefront finishes;
//   * primaries are injected from the resident source records whenever the
//     number of particles in flight drops below a refill threshold;
//   * a sweep injects primaries, then runs its photon side (the photon
//     kernel, then one kernel for the four photon interactions) and its
//     electron side (the electron kernel, then one kernel for the four
//     electron interactions) concurrently on two streams, each side draining
//     what the previous sweeps left for it, then does its bookkeeping;
//   * a group of sweeps is captured in a hipGraph and replayed until no
//     particle and no source record is left, or until every record is
//     injected and few particles remain, which one tail launch then
//     finishes, each particle with its descendants in one thread; the host
//     reads the loop's state once per replay.
// Particles keep their Philox streams across banks, so every history follows
// exactly the trajectory the plain kernel and the oracle give it.
#pragma once

// Marks a function whose wavefront operations are shared by exactly the
// lanes that call it together -- a subset of the wavefront when the call is
// in divergent code -- unlike the kernels' other wavefront operations, which
// every lane of the wavefront reaches together.  Nothing on the GPU; the CPU
// build with 64-lane wavefronts groups lanes by it.
#ifndef GEGS_WAVE_SUBSET_OP
#define GEGS_WAVE_SUBSET_OP(id)
#endif
#include <vector>

#include "plain.hpp"
#include "transport.hpp"

namespace gegs {

constexpr int EVW = 21;        // words per banked particle (15-19: state of a suspended electron step, 20: job)
constexpr int SLOT_P_ALT = N_EVENTS;         // second buffer of the photon-step bank
constexpr int SLOT_E_ALT = N_EVENTS + 1;     // second buffer of the electron-step bank
constexpr int N_SLOTS = N_EVENTS + 2;

// A queue counter alone in a 256-byte block.  Device-scope atomics are
// performed beyond the L2 of the XCD that issues them; while one is out, its
// L2 channel holds the line, and the channel's later requests wait behind
// it.  Counters that share a line serialize each other's atomics, and a
// wavefront load with any lane mapped to that channel (most gathers of 64
// scattered voxel classes have one) waits with them.
struct alignas(256) QCount {
    uint32_t v;
};

// Every bank is split into QSEG equal append segments, each with its own
// counter: block b appends to segment b % QSEG.  Blocks are dealt to the 8
// XCDs in turn, so a segment's counter is used by one XCD, and each XCD
// spreads its appends to one bank over QSEG / 8 lines instead of queueing
// them on one.  Consumers see a bank as one range through BankView.
constexpr int QSEG = 32;

struct EvState {
    uint32_t overflow;               // particles that did not fit
    uint32_t active;                 // more work remains after the sweep (read by the host with `sweeps`)
    uint32_t sweeps;
    uint32_t flip;                   // buffer of banks 0 and 5 drained this sweep
    uint32_t src_next, src_last;     // source records still to inject
    uint32_t inj_start, inj_n;       // current injection
    uint32_t inflight;               // particles in the banks when the last sweep ended
    QCount count[N_SLOTS][QSEG];     // particles appended to each segment of each bank
    QCount taken[N_SLOTS];           // work handed out from each bank (over its whole range)
};

struct EventConfig {
    uint32_t block = 256;            // threads per block of the single-pass event kernels (<= 256)
    uint32_t flight_block = 256;     // threads per block of the photon and electron kernels (<= 512)
    uint32_t grid = 2048;            // blocks of the single-pass interaction kernels
    uint32_t photon_grid = 0;        // blocks of the photon kernel (0: fill the device once)
    uint32_t electron_grid = 0;      // blocks of the electron kernel (0: fill the device once)
    uint32_t photon_chunk = 32;      // voxel crossings per photon and sweep (0: whole flight)
    uint32_t electron_chunk = 16;    // condensed-history steps per electron and sweep (0: no limit)
    uint32_t refill_min = 16;        // idle lanes of a wavefront that trigger a refill
    uint32_t refill_slots = 64;      // bank slots a wavefront reserves per refill atomic
    uint32_t sweep_streams = 2;      // 2: photon and electron sides of a sweep concurrently; 1: one after the other
    uint32_t tail = 1u << 18;        // in-flight particles that one tail launch finishes (0: no tail launch)
    size_t capacity = 0;             // slots per bank (0: twice what can be in flight, at least 4 Mi)
    uint32_t inject = 6u << 20;      // primaries injected per refill
    uint32_t threshold = 6u << 20;   // refill when fewer particles are in flight
    uint32_t sweeps_per_graph = 4;
};

// Source records of one batch, as slices of every job (field) of the run:
// cursor k of the batch belongs to job j with prefix[j] <= k < prefix[j+1];
// its record is rec0[j] + (k - prefix[j]), its Philox history index
// hist0[j] + (k - prefix[j]) and its per-history tally slot slot0[j] + (k - prefix[j]).
constexpr uint32_t MAX_JOBS = 512;

struct BatchPlan {
    uint32_t njob;
    uint32_t total;                  // histories in the batch
    uint32_t prefix[MAX_JOBS + 1];
    uint32_t rec0[MAX_JOBS];
    uint32_t slot0[MAX_JOBS];
    uint64_t hist0[MAX_JOBS];
};

struct EvArgs {
    Env env;
    const BatchPlan *plan;
    const uint32_t *job_plan;        // dose buffer (plan) of each job
    size_t nvox;
    float *bank;                     // [slot][word][cap]
    uint32_t cap;
    uint32_t seg_cap;                // cap / QSEG: slots per append segment
    EvState *st;
    const gegs_source *src;
    double *dose;
    double *hist_edep;               // per source record, optional
    uint32_t *hist_steps;            // optional
    unsigned long long *counters;
    // beamlet weights: history slot k belongs to beamlet beamlet[k]; the
    // forward multiplies its source weight by bweight[beamlet[k]]
    const uint32_t *beamlet;         // optional
    const float *bweight;            // optional
    // adjoint replay: dL/dD per plan and voxel, gradient per beamlet
    const double *adj;
    double *grad;
    // statistical batches tallied in one pass: a deposit of history slot h
    // of job j goes to dose row (plan(j) * tally_batches + batch(h)); 1: one
    // row per plan
    uint32_t tally_batches;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// the lowest active lane issues a wavefront's shared atomics (lane 0 where
// the whole wavefront is active)
__device__ __forceinline__ bool first_lane() { return (__ballot(1) & lanes_below()) == 0; }

// A bank's particles as one range [0, total()): segment g contributes
// [pre[g], pre[g+1]), stored at bank slots g * seg_cap + [0, pre[g+1] - pre[g]).
// Counts past seg_cap are appends that overflowed (counted in `overflow`).
struct BankView {
    uint32_t pre[QSEG + 1];
    __device__ __forceinline__ void load(const EvArgs &a, int slot) {
        pre[0] = 0;
#pragma unroll
        for (int g = 0; g < QSEG; ++g) {
            const uint32_t c = a.st->count[slot][g].v;
            pre[g + 1] = pre[g] + (c < a.seg_cap ? c : a.seg_cap);
        }
    }
    __device__ __forceinline__ uint32_t total() const { return pre[QSEG]; }
    // bank slot of range position k < total(): the last segment that starts
    // at or before k holds it (empty segments start where the next one does)
    __device__ __forceinline__ uint32_t slot_of(const EvArgs &a, uint32_t k) const {
        uint32_t base = 0, off = k;
#pragma unroll
        for (int g = 1; g < QSEG; ++g) {
            if (k >= pre[g]) {
                base = (uint32_t)g * a.seg_cap;
                off = k - pre[g];
            }
        }
        return base + off;
    }
};

// bank buffer of event e drained in this sweep (now) or in the next one
__device__ __forceinline__ int bank_slot(int e, uint32_t flip, bool now) {
    const bool alt = (flip != 0) == now;
    if (e == P_PHOTON_STEP) return alt ? SLOT_P_ALT : P_PHOTON_STEP;
    if (e == E_ELECTRON_STEP) return alt ? SLOT_E_ALT : E_ELECTRON_STEP;
    return e;
}

__device__ __forceinline__ float *bank_word(const EvArgs &a, int slot, int w, uint32_t idx) {
    return a.bank + ((size_t)slot * EVW + w) * a.cap + idx;
}

// Bank records stream through once per event hop: non-temporal loads and
// stores keep them from evicting the voxel class volume and the tables that
// every crossing reads from L2.
__device__ __forceinline__ void bank_put(const EvArgs &a, int slot, int w, uint32_t idx, float v) {
    __builtin_nontemporal_store(v, bank_word(a, slot, w, idx));
}

__device__ __forceinline__ float bank_get(const EvArgs &a, int slot, int w, uint32_t idx) {
    return __builtin_nontemporal_load(bank_word(a, slot, w, idx));
}

__device__ __forceinline__ void bank_store(const EvArgs &a, int slot, uint32_t idx, const Part &p, uint32_t hist) {
    bank_put(a, slot, 0, idx, p.e);
    bank_put(a, slot, 1, idx, p.x);
    bank_put(a, slot, 2, idx, p.y);
    bank_put(a, slot, 3, idx, p.z);
    bank_put(a, slot, 4, idx, p.u);
    bank_put(a, slot, 5, idx, p.v);
    bank_put(a, slot, 6, idx, p.w);
    bank_put(a, slot, 7, idx, p.wt);
    bank_put(a, slot, 8, idx, p.dnear);
    bank_put(a, slot, 9, idx, __uint_as_float(p.region));
    bank_put(a, slot, 10, idx,
             __uint_as_float((uint32_t)(p.charge + 1) | ((uint32_t)p.spos << 8) | (((uint32_t)p.module & 7u) << 16) |
                             (((uint32_t)(p.bank + 1) & 7u) << 19) | ((uint32_t)(p.leaf + 128) << 24)));
    bank_put(a, slot, 11, idx, __uint_as_float((uint32_t)p.lin));
    bank_put(a, slot, 12, idx, __uint_as_float((uint32_t)(p.lin >> 32)));
    bank_put(a, slot, 13, idx, __uint_as_float(p.blk));
    bank_put(a, slot, 14, idx, __uint_as_float(hist));
    bank_put(a, slot, 20, idx, __uint_as_float(p.job));
}

__device__ __forceinline__ Part bank_load(const EvArgs &a, int slot, int event, uint32_t idx, uint32_t &hist) {
    Part p;
    p.e = bank_get(a, slot, 0, idx);
    p.x = bank_get(a, slot, 1, idx);
    p.y = bank_get(a, slot, 2, idx);
    p.z = bank_get(a, slot, 3, idx);
    p.u = bank_get(a, slot, 4, idx);
    p.v = bank_get(a, slot, 5, idx);
    p.w = bank_get(a, slot, 6, idx);
    p.wt = bank_get(a, slot, 7, idx);
    p.dnear = bank_get(a, slot, 8, idx);
    p.region = __float_as_uint(bank_get(a, slot, 9, idx));
    const uint32_t pk = __float_as_uint(bank_get(a, slot, 10, idx));
    p.charge = (int32_t)(pk & 0xffu) - 1;
    p.spos = (int32_t)((pk >> 8) & 15u);
    p.module = (int32_t)((pk >> 16) & 7u);
    p.bank = (int32_t)((pk >> 19) & 7u) - 1;
    p.leaf = (int32_t)(pk >> 24) - 128;
    p.status = event;
    p.lin = (uint64_t)__float_as_uint(bank_get(a, slot, 11, idx)) |
            ((uint64_t)__float_as_uint(bank_get(a, slot, 12, idx)) << 32);
    p.blk = __float_as_uint(bank_get(a, slot, 13, idx));
    hist = __float_as_uint(bank_get(a, slot, 14, idx));
    p.job = __float_as_uint(bank_get(a, slot, 20, idx));
    return p;
}

// deposits of a particle go to the dose buffer (forward) or are weighted by
// the dL/dD (adjoint) of its job's plan
// The batch of slot `hist` is the b-th of tally_batches equal slices of its
// job, the slices of sequential batches: [floor(count b / nb), floor(count (b+1) / nb)).
template <bool ADJ>
__device__ __forceinline__ void bind_tally(const EvArgs &a, const Part &p, uint32_t hist, TallyT<ADJ> &t) {
    size_t row = a.job_plan[p.job];
    if (!ADJ && a.tally_batches > 1) {
        const BatchPlan &bp = *a.plan;
        const uint64_t count = bp.prefix[p.job + 1] - bp.prefix[p.job];
        const uint64_t off = hist - bp.slot0[p.job];
        const uint64_t nb = a.tally_batches;
        row = row * nb + ((off + 1) * nb - 1) / count;
    }
    const size_t off = row * a.nvox;
    if (ADJ) t.adj = a.adj + off;
    else t.dose = a.dose + off;
}

// Appends to the bank named by p.status.  The lanes that reach a push
// together are grouped by destination; each group reserves its slots with
// one atomic issued by its first lane.  Photon-step and electron-step pushes
// go to the buffer drained in this sweep when their kernel has not started
// in the sweep (photons_now / electrons_now), else to the next sweep's.  The
// photon side of a sweep (photon kernel, photon interactions) and its
// electron side run concurrently, so only the injection, which precedes
// both, pushes to the current buffers.
struct BankSink {
    const EvArgs *a;
    uint32_t hist;
    uint32_t flip;
    bool photons_now, electrons_now;
    // The groups are served one per iteration by every lane of the call:
    // `todo` (a wavefront-uniform mask) loses the served group each time and
    // the loop ends when it is empty.  A loop whose only exit is a divergent
    // `dest == readfirstlane(dest)` test must not be used here: its state does
    // not change between iterations, so under C++ forward-progress rules the
    // compiler may assume the first iteration exits and drop the test, which
    // sends every lane's particle to the first lane's destination count.
    __device__ __forceinline__ void push(const Part &p) {
        GEGS_WAVE_SUBSET_OP(1);
        const uint32_t seg = blockIdx.x % QSEG;
        const int e = p.status;
        const int dest = bank_slot(e, flip, e == P_PHOTON_STEP ? photons_now : electrons_now);
        uint64_t todo = __ballot(1);
        while (todo) {
            const int leader = __builtin_ctzll(todo);
            const int d0 = __builtin_amdgcn_readlane(dest, leader);
            const bool mine = dest == d0;
            const uint64_t m = __ballot(mine);
            uint32_t base = 0;
            if ((int)lane_id() == leader) base = atomicAdd(&a->st->count[d0][seg].v, (uint32_t)__popcll(m));
            base = __builtin_amdgcn_readlane(base, leader);
            if (mine) {
                const uint32_t idx = base + (uint32_t)__popcll(m & lanes_below());
                if (idx < a->seg_cap) bank_store(*a, d0, seg * a->seg_cap + idx, p, hist);
                else atomicAdd(&a->st->overflow, 1u);
            }
            todo &= ~m;
        }
    }
    // an electron suspended between two condensed-history steps carries the
    // step state (electron_iter's EState) to the sweep that resumes it
    __device__ __forceinline__ void push_suspended(const Part &p, const EState &es) {
        GEGS_WAVE_SUBSET_OP(2);
        const uint32_t seg = blockIdx.x % QSEG;
        const int dest = bank_slot(p.status, flip, electrons_now);
        const uint64_t m = __ballot(1);
        const uint32_t rank = (uint32_t)__popcll(m & lanes_below());
        uint32_t base = 0;
        if (rank == 0) base = atomicAdd(&a->st->count[dest][seg].v, (uint32_t)__popcll(m));
        base = __builtin_amdgcn_readfirstlane(base);
        const uint32_t at = base + rank;
        const uint32_t idx = seg * a->seg_cap + at;
        if (at < a->seg_cap) {
            bank_store(*a, dest, idx, p, hist);
            bank_put(*a, dest, 15, idx, es.elke);
            bank_put(*a, dest, 16, idx, es.demfp);
            bank_put(*a, dest, 17, idx, es.sig0);
            bank_put(*a, dest, 18, idx, es.total_tstep);
            bank_put(*a, dest, 19, idx,
                     __uint_as_float((uint32_t)es.lelke | ((uint32_t)es.compute_tstep << 16) |
                                     ((uint32_t)es.is_tstep << 17)));
        } else {
            atomicAdd(&a->st->overflow, 1u);
        }
    }
};

__device__ __forceinline__ void load_estate(const EvArgs &a, int slot, uint32_t idx, EState &es) {
    es.elke = bank_get(a, slot, 15, idx);
    es.demfp = bank_get(a, slot, 16, idx);
    es.sig0 = bank_get(a, slot, 17, idx);
    es.total_tstep = bank_get(a, slot, 18, idx);
    const uint32_t f = __float_as_uint(bank_get(a, slot, 19, idx));
    es.lelke = (int)(f & 0xffffu);
    es.compute_tstep = (f >> 16) & 1u;
    es.is_tstep = (f >> 17) & 1u;
}

template <bool ADJ>
__device__ __forceinline__ void flush_particle(const EvArgs &a, uint32_t hist, TallyT<ADJ> &t) {
    if (ADJ) {
        if (t.acc != 0.0) unsafeAtomicAdd(a.grad + a.beamlet[hist], t.acc);
        t.acc = 0.0;
    }
    if (a.hist_edep && t.hist_edep != 0.0) unsafeAtomicAdd(a.hist_edep + hist, t.hist_edep);
    if (a.hist_steps && t.steps) atomicAdd(a.hist_steps + hist, t.steps);
    t.hist_edep = 0.0;
    t.steps = 0;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// counters are reduced over the wavefront first: one atomic per wave and
// counter instead of one per lane (all lanes of the wave must call this)
template <class T>
__device__ __forceinline__ void flush_counters(const EvArgs &a, const T &t, int event, uint32_t pops) {
    unsigned long long *c = a.counters;
    const uint32_t v0 = wave_sum(t.photon_steps), v1 = wave_sum(t.electron_steps);
    const uint32_t v2 = wave_sum(t.gathers), v3 = wave_sum(t.deposits), v4 = wave_sum(pops);
    if (first_lane()) {
        if (v0) atomicAdd(c + 0, (unsigned long long)v0);
        if (v1) atomicAdd(c + 1, (unsigned long long)v1);
        if (v2) atomicAdd(c + 2, (unsigned long long)v2);
        if (v3) atomicAdd(c + 3, (unsigned long long)v3);
        if (v4) atomicAdd(c + 4 + event, (unsigned long long)v4);
    }
}

template <bool ADJ>
__device__ __forceinline__ TallyT<ADJ> fresh_tally(const EvArgs &a) {
    TallyT<ADJ> t;
    t.dose = a.dose;
    t.adj = a.adj;
    t.acc = 0.0;
    t.hist_edep = 0.0;
    t.steps = 0;
    t.photon_steps = t.electron_steps = t.gathers = t.deposits = 0;
    return t;
}

// One particle of single-pass event E: the interaction's sampler.
template <int E, bool ADJ>
__device__ __forceinline__ void interact(const EvArgs &a, Stream &s, Part &p, BankSink &sink, TallyT<ADJ> &t) {
    if (E == P_COMPTON) compton(a.env, s, p, sink, t);
    else if (E == P_PHOTO) photo(a.env, s, p, sink, t);
    else if (E == P_RAYLEIGH) rayleigh(a.env, s, p, sink, t);
    else if (E == P_PAIR) pair(a.env, s, p, sink, t);
    else if (E == E_MOLLER) moller(a.env, s, p, sink, t);
    else if (E == E_BREMS) brems(a.env, s, p, sink, t);
    else if (E == E_BHABHA) bhabha(a.env, s, p, sink, t);
    else annih(a.env, s, p, sink, t);
}

// particles in a bank (what BankView::total gives, without the prefix)
__device__ __forceinline__ uint32_t bank_total(const EvArgs &a, int slot) {
    uint32_t n = 0;
#pragma unroll
    for (int g = 0; g < QSEG; ++g) {
        const uint32_t c = a.st->count[slot][g].v;
        n += c < a.seg_cap ? c : a.seg_cap;
    }
    return n;
}

// The four single-pass events of one particle type in one launch: E0 = 1
// (compton, photo, rayleigh, pair) or E0 = 6 (moller, brems, bhabha, annih).
// The banks' particles are dealt in chunks of 64, bank after bank, and
// wavefront w takes chunks w, w + nwaves, ...; a chunk holds particles of one
// bank only, so a wavefront runs one sampler at a time.  Interactions create
// particles only in the photon-step and electron-step banks, so the four
// banks' counts hold for the whole launch.  One launch instead of four
// removes three kernel boundaries from every sweep; in the last sweeps of a
// run, when the banks hold a few particles each, the boundaries are most of
// the time these events take.
template <int E0, bool ADJ>
__global__ void __launch_bounds__(256) interaction_kernel(EvArgs a) {
    static_assert(E0 == P_COMPTON || E0 == E_MOLLER, "a group is the four single-pass events of one type");
    uint32_t c1, c2, c3, c4;                 // chunks of the banks before bank E0 + 1, + 2, + 3, and in all
    c1 = (bank_total(a, E0) + 63u) >> 6;
    c2 = c1 + ((bank_total(a, E0 + 1) + 63u) >> 6);
    c3 = c2 + ((bank_total(a, E0 + 2) + 63u) >> 6);
    c4 = c3 + ((bank_total(a, E0 + 3) + 63u) >> 6);
    if (c4 == 0) return;
    TallyT<ADJ> t = fresh_tally<ADJ>(a);
    BankSink sink{&a, 0, a.st->flip, false, false};
    uint32_t pops0 = 0, pops1 = 0, pops2 = 0, pops3 = 0;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    BankView view;
    int cur = -1;
    uint32_t n = 0;
    for (uint32_t c = wave; c < c4; c += nwaves) {
        const int k = (c >= c1) + (c >= c2) + (c >= c3);     // wave-uniform
        const uint32_t first = k == 0 ? 0u : (k == 1 ? c1 : (k == 2 ? c2 : c3));
        if (k != cur) {
            view.load(a, E0 + k);
            n = view.total();
            cur = k;
        }
        const uint32_t idx = (c - first) * 64u + lane_id();
        if (idx < n) {
            uint32_t hist;
            Part p = bank_load(a, E0 + k, E0 + k, view.slot_of(a, idx), hist);
            bind_tally(a, p, hist, t);
            sink.hist = hist;
            Stream s;
            s.begin(p, a.env.seed);
            if (k == 0) {
                interact<E0, ADJ>(a, s, p, sink, t);
                pops0++;
            } else if (k == 1) {
                interact<E0 + 1, ADJ>(a, s, p, sink, t);
                pops1++;
            } else if (k == 2) {
                interact<E0 + 2, ADJ>(a, s, p, sink, t);
                pops2++;
            } else {
                interact<E0 + 3, ADJ>(a, s, p, sink, t);
                pops3++;
            }
            flush_particle(a, hist, t);
        }
    }
    flush_counters(a, t, E0, pops0);
    const uint32_t v1 = wave_sum(pops1), v2 = wave_sum(pops2), v3 = wave_sum(pops3);
    if (first_lane()) {
        if (v1) atomicAdd(a.counters + 4 + E0 + 1, (unsigned long long)v1);
        if (v2) atomicAdd(a.counters + 4 + E0 + 2, (unsigned long long)v2);
        if (v3) atomicAdd(a.counters + 4 + E0 + 3, (unsigned long long)v3);
    }
}

// Lane refill for the looping events: a wavefront reserves `slots` (>= 64)
// bank slots with one atomic and hands them to its idle lanes in order,
// reserving again when they run out.  Every reserved slot is transported by
// the wavefront that holds it: the loop ends only when the reservation is
// used up and the bank is exhausted.
struct Refill {
    uint32_t next, end;      // wave-uniform reservation
    bool exhausted;
    __device__ __forceinline__ void init() {
        next = end = 0;
        exhausted = false;
    }
    // returns the slot for this lane (>= n: none); call with all lanes active.
    // Idle lanes are refilled together once at least min_idle of them wait
    // (or the whole wavefront is idle), so that the loads that start a
    // particle are paid once for many lanes rather than on every iteration.
    __device__ __forceinline__ uint32_t take(uint32_t *taken, uint32_t n, bool want, uint32_t min_idle,
                                             uint32_t slots) {
        const uint64_t idle = __ballot(want);
        uint32_t need = (uint32_t)__popcll(idle);
        uint32_t slot = 0xFFFFFFFFu;
        const uint32_t rank = (uint32_t)__popcll(idle & lanes_below());
        if (need == 0 || exhausted || (need < min_idle && need < 64u)) return slot;
        if (end - next < need) {
            // top up the reservation: keep the remainder, add `slots` more
            uint32_t got = 0;
            if (first_lane()) got = atomicAdd(taken, slots);
            got = __builtin_amdgcn_readfirstlane(got);
            if (got >= n) {
                exhausted = true;
            } else {
                // remainder [next, end) is used first, then [got, got + slots)
                const uint32_t rem = end - next;
                if (want && rank < rem) slot = next + rank;
                else if (want) slot = got + (rank - rem);
                const uint32_t used_new = need > rem ? need - rem : 0u;
                next = got + used_new;
                end = got + slots;
                return slot < n ? slot : 0xFFFFFFFFu;
            }
        }
        if (want && rank < end - next) slot = next + rank;
        const uint32_t used = need < end - next ? need : end - next;
        next += used;
        return slot < n ? slot : 0xFFFFFFFFu;
    }
};

// A flight ends with at most one push, of the photon itself (to the bank of
// its interaction, or suspended): the walk marks it here and the particle
// stays in the lane's registers until the wavefront next refills.
struct HoldSink {
    bool held;
    __device__ __forceinline__ void push(const Part &) { held = true; }
};

// photon free flight with lane refill; a flight still going after `chunk`
// crossings is suspended into the next sweep's bank.  The photons of ended
// flights are appended together when the wavefront refills (or runs dry),
// so that one atomic per destination serves many lanes.
template <bool TAB, bool ADJ, bool GEN>
__global__ void __launch_bounds__(512) photon_kernel(EvArgs a, uint32_t chunk, uint32_t refill_min,
                                                     uint32_t refill_slots) {
    extern __shared__ float lds_env[];
    const uint32_t flip = a.st->flip;
    const int slot_in = bank_slot(P_PHOTON_STEP, flip, true);
    BankView view;
    view.load(a, slot_in);
    const uint32_t n = view.total();
    if (n == 0) return;
    Env env = a.env;
    env.fixes = kernel_fixes<GEN>(env.fixes);
    stage_env<TAB>(env, lds_env);
    TallyT<ADJ> t = fresh_tally<ADJ>(a);
    BankSink sink{&a, 0, flip, false, false};
    HoldSink hold{false};
    uint32_t pops = 0;
    bool have = false, interact = false;
    Part p;
    Stream s;
    PhotonWalk w;
    uint32_t hist = 0, iters = 0;
    Refill rf;
    rf.init();
    for (;;) {
        // the held photons go out before a refill can overwrite their lanes:
        // the same condition as in Refill::take, or no refill is coming
        const uint32_t nidle = (uint32_t)__popcll(__ballot(!have));
        if (nidle && (nidle >= refill_min || nidle == 64u || rf.exhausted) && __any(hold.held)) {
            if (hold.held) {
                // flights that reached their interaction point choose it now,
                // together, from their untouched streams
                if (interact) {
                    photon_interaction(env, s, p, w, hold);
                    interact = false;
                }
                sink.hist = hist;
                sink.push(p);
                hold.held = false;
            }
        }
        const uint32_t slot = rf.take(&a.st->taken[slot_in].v, n, !have, refill_min, refill_slots);
        if (!have && slot != 0xFFFFFFFFu) {
            p = bank_load(a, slot_in, P_PHOTON_STEP, view.slot_of(a, slot), hist);
            bind_tally(a, p, hist, t);
            s.begin(p, env.seed);
            photon_walk_begin(env, p, w);
            have = true;
            iters = 0;
            pops += (p.spos & 8) ? 0u : 1u;    // flights, not resumptions
        }
        if (!__any(have)) break;
        wave_lookahead(s, have);
        if (have) {
            if (chunk && iters == chunk) {
                suspend(hold, s, p, P_PHOTON_STEP);
                have = false;
                flush_particle(a, hist, t);
            } else {
                const int r = photon_walk_iter(env, s, p, w, hold, t);
                if (r != PW_ON) {
                    have = false;
                    flush_particle(a, hist, t);
                    if (r == PW_INTERACT) {
                        interact = true;
                        hold.held = true;
                    }
                }
            }
            iters++;
        }
    }
    flush_counters(a, t, P_PHOTON_STEP, pops);
}

// electron / positron condensed-history steps with lane refill; an electron
// still stepping after `chunk` steps is suspended into the next sweep's bank
template <bool TAB, bool ADJ, bool GEN>
__global__ void __launch_bounds__(512) GEGS_WAVES_PER_EU(3) electron_kernel(EvArgs a, uint32_t chunk, uint32_t refill_min,
                                                       uint32_t refill_slots) {
    extern __shared__ float lds_env[];
    const uint32_t flip = a.st->flip;
    const int slot_in = bank_slot(E_ELECTRON_STEP, flip, true);
    BankView view;
    view.load(a, slot_in);
    const uint32_t n = view.total();
    if (n == 0) return;
    Env env = a.env;
    env.fixes = kernel_fixes<GEN>(env.fixes);
    stage_env<TAB>(env, lds_env);
    TallyT<ADJ> t = fresh_tally<ADJ>(a);
    BankSink sink{&a, 0, flip, false, false};
    uint32_t pops = 0;
    bool have = false;
    Part p;
    Stream s;
    EState es;
    uint32_t hist = 0, iters = 0;
    Refill rf;
    rf.init();
    for (;;) {
        const uint32_t slot = rf.take(&a.st->taken[slot_in].v, n, !have, refill_min, refill_slots);
        if (!have && slot != 0xFFFFFFFFu) {
            const uint32_t at = view.slot_of(a, slot);
            p = bank_load(a, slot_in, E_ELECTRON_STEP, at, hist);
            bind_tally(a, p, hist, t);
            s.begin(p, env.seed);
            if (p.spos & 8) {
                load_estate(a, slot_in, at, es);
            } else {
                es.init();
                pops++;
            }
            have = true;
            iters = 0;
        }
        if (!__any(have)) break;
        wave_lookahead_eager(s, have);
        if (have) {
            sink.hist = hist;
            if (chunk && iters == chunk) {
                p.status = E_ELECTRON_STEP;
                p.lin = s.lin;
                p.blk = s.blk;
                p.spos = s.pos | 8;
                sink.push_suspended(p, es);
                have = false;
                flush_particle(a, hist, t);
            } else if (!electron_iter(env, s, p, es, sink, t)) {
                have = false;
                flush_particle(a, hist, t);
            }
            iters++;
        }
    }
    flush_counters(a, t, E_ELECTRON_STEP, pops);
}

// plan the next sweep's injection: refill when few particles are in flight
__device__ __forceinline__ void plan_injection(EvState *st, uint32_t threshold, uint32_t max_inject) {
    uint32_t n = 0;
    if (st->inflight < threshold && st->src_next < st->src_last) {
        n = st->src_last - st->src_next;
        if (n > max_inject) n = max_inject;
    }
    st->inj_start = st->src_next;
    st->inj_n = n;
    st->src_next += n;
}

__global__ void __launch_bounds__(256) inject_fill(EvArgs a) {
    const uint32_t n = a.st->inj_n;
    const uint32_t start = a.st->inj_start;
    BankSink sink{&a, 0, a.st->flip, true, true};
    const BatchPlan &bp = *a.plan;
    const uint32_t njob = bp.njob;
    uint32_t injected = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    // whole wavefronts iterate together so that the aggregated push sees
    // the lanes that have a record in the same pass
    const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
    for (uint32_t k0 = wave0; k0 < n; k0 += stride) {
        const uint32_t k = k0 + lane_id();
        if (k < n) {
            const uint32_t cur = start + k;
            // job of the cursor: last j with prefix[j] <= cur
            uint32_t lo = 0, hi = njob;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (bp.prefix[mid] <= cur) lo = mid; else hi = mid;
            }
            const uint32_t off = cur - bp.prefix[lo];
            const uint32_t rec = bp.rec0[lo] + off;
            const gegs_source r = a.src[rec];
            Part p;
            p.e = r.e; p.x = r.x; p.y = r.y; p.z = r.z;
            p.u = r.u; p.v = r.v; p.w = r.w; p.wt = r.wt;
            p.charge = r.charge;
            p.region = 0;
            p.job = lo;
            p.module = a.env.fields[lo].module_first;
            p.leaf = OUTMLC;
            p.bank = -1;
            p.dnear = INFI_DIST;
            p.blk = 0;
            p.spos = 4;
            p.lin = primary_lineage(bp.hist0[lo] + off);
            p.status = r.charge == 0 ? P_PHOTON_STEP : E_ELECTRON_STEP;
            sink.hist = bp.slot0[lo] + off;
            if (a.bweight) p.wt = p.wt * a.bweight[a.beamlet[sink.hist]];
            sink.push(p);
            injected++;
        }
    }
    const uint32_t tot = wave_sum(injected);
    if (first_lane() && tot) atomicAdd(a.counters + 14, (unsigned long long)tot);
}

__global__ void sweep_end(EvState *st, uint32_t threshold, uint32_t max_inject) {
    // every bank drained in this sweep is empty: the eight single-pass banks
    // and the current buffers of banks 0 and 5, which take the next sweep's
    // pushes after the flip
    const int p_in = st->flip ? SLOT_P_ALT : P_PHOTON_STEP;
    const int e_in = st->flip ? SLOT_E_ALT : E_ELECTRON_STEP;
    const int drained[10] = {P_COMPTON, P_PHOTO, P_RAYLEIGH, P_PAIR, E_MOLLER, E_BREMS, E_BHABHA, E_ANNIH, p_in, e_in};
    for (int k = 0; k < 10; ++k) {
        for (int g = 0; g < QSEG; ++g) st->count[drained[k]][g].v = 0;
        st->taken[drained[k]].v = 0;
    }
    // what is left in flight is in the other buffers of banks 0 and 5
    const int p_next = st->flip ? P_PHOTON_STEP : SLOT_P_ALT;
    const int e_next = st->flip ? E_ELECTRON_STEP : SLOT_E_ALT;
    uint32_t inflight = 0;
    for (int g = 0; g < QSEG; ++g) inflight += st->count[p_next][g].v + st->count[e_next][g].v;
    st->flip ^= 1u;
    st->inflight = inflight;
    st->active = (inflight > 0 || st->src_next < st->src_last) ? 1u : 0u;
    st->sweeps += 1;
    plan_injection(st, threshold, max_inject);
}

// The last particles of a run, each with all its descendants in one thread.
// Once every source record is injected and at most `tail` particles are in
// flight, the sweeps would go on one interaction or one chunk of crossings
// at a time, each sweep paying its launches and bank round trips for a few
// particles; this launch takes the particles the last sweep left (photon
// and electron banks) and transports each to the end of its history branch
// as the plain kernel does, secondaries on a per-thread stack, interactions
// in line.  A particle's trajectory does not depend on which kernel moves
// it (each carries its stream), so the histories are the sweeps' own; a
// suspended electron resumes from its banked step state.
template <bool TAB, bool ADJ, bool GEN>
__global__ void __launch_bounds__(128) GEGS_WAVES_PER_EU(2) tail_kernel(EvArgs a, float *stack, int depth) {
    extern __shared__ float lds_env[];
    const uint32_t flip = a.st->flip;
    const int sp_in = bank_slot(P_PHOTON_STEP, flip, true);
    const int se_in = bank_slot(E_ELECTRON_STEP, flip, true);
    BankView vp, ve;
    vp.load(a, sp_in);
    ve.load(a, se_in);
    const uint32_t np = vp.total(), n = np + ve.total();
    if (n == 0) return;
    Env env = a.env;
    env.fixes = kernel_fixes<GEN>(env.fixes);
    stage_env<TAB>(env, lds_env);
    TallyT<ADJ> t = fresh_tally<ADJ>(a);
    StackSink sink;
    sink.base = stack;
    sink.stride = gridDim.x * blockDim.x;
    sink.tid = blockIdx.x * blockDim.x + threadIdx.x;
    sink.depth = depth;
    sink.overflow = 0;
    sink.sp = 0;
    uint32_t pops[N_EVENTS];
#pragma unroll
    for (int k = 0; k < N_EVENTS; ++k) pops[k] = 0;
    for (uint32_t k = sink.tid; k < n; k += sink.stride) {
        const bool photon = k < np;
        const uint32_t at = photon ? vp.slot_of(a, k) : ve.slot_of(a, k - np);
        uint32_t hist;
        Part p = photon ? bank_load(a, sp_in, P_PHOTON_STEP, at, hist) : bank_load(a, se_in, E_ELECTRON_STEP, at, hist);
        bind_tally(a, p, hist, t);
        Stream s;
        s.begin(p, env.seed);
        sink.sp = 0;
        const bool fresh = (p.spos & 8) == 0;     // a new flight or track, not a resumed one
        if (photon) {
            pops[P_PHOTON_STEP] += fresh ? 1u : 0u;
            photon_step<false>(env, s, p, sink, t);
        } else {
            EState es;
            if (fresh) es.init();
            else load_estate(a, se_in, at, es);
            pops[E_ELECTRON_STEP] += fresh ? 1u : 0u;
            while (electron_iter(env, s, p, es, sink, t)) {
            }
        }
        while (sink.sp > 0) {
            Part q = sink.pop();
            Stream sq;
            sq.begin(q, env.seed);
#pragma unroll
            for (int e = 0; e < N_EVENTS; ++e) pops[e] += (q.status == e);
            handle<false>(env, sq, q, sink, t);
        }
        flush_particle(a, hist, t);
    }
    if (sink.overflow) atomicAdd(&a.st->overflow, sink.overflow);
    flush_counters(a, t, P_PHOTON_STEP, pops[P_PHOTON_STEP]);
    unsigned long long *c = a.counters;
#pragma unroll
    for (int e = 1; e < N_EVENTS; ++e) {
        const uint32_t v = wave_sum(pops[e]);
        if (first_lane() && v) atomicAdd(c + 4 + e, (unsigned long long)v);
    }
}

// after the tail launch: every bank is empty and nothing is left to inject
__global__ void tail_end(EvState *st) {
    for (int e = 0; e < N_SLOTS; ++e) {
        for (int g = 0; g < QSEG; ++g) st->count[e][g].v = 0;
        st->taken[e].v = 0;
    }
    st->inflight = 0;
    st->active = 0;
}

__global__ void batch_begin(EvState *st, uint32_t first, uint32_t last, uint32_t threshold, uint32_t max_inject) {
    for (int e = 0; e < N_SLOTS; ++e) {
        for (int g = 0; g < QSEG; ++g) st->count[e][g].v = 0;
        st->taken[e].v = 0;
    }
    st->inflight = 0;
    st->flip = 0;
    st->src_next = first;
    st->src_last = last;
    st->inj_start = first;
    st->inj_n = 0;
    st->active = 1;
    st->sweeps = 0;
    plan_injection(st, threshold, max_inject);
}

__global__ void clear_overflow(EvState *st) { st->overflow = 0; }

__global__ void report_overflow(const EvState *st, unsigned long long *counters) {
    counters[15] += (unsigned long long)st->overflow;
}

// ------------------------------------------------------------------ host side
// EvState words the host reads after each replay: active, sweeps, flip,
// src_next, src_last, inj_start, inj_n, inflight
constexpr int FLAG_WORDS = 8;
// The second stream of a sweep: the electron side runs on `side` between
// the events fork (recorded on the sweep's stream after the injection) and
// join (waited for before the sweep's bookkeeping).  side == nullptr: the
// whole sweep on one stream, photon side first.
struct SweepFork {
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

struct EventBanks {
    float *bank = nullptr;
    uint32_t cap = 0;
    EvState *st = nullptr;
    uint32_t *h_flags = nullptr;     // pinned, per replay slot: EvState from `active` to `inflight`
    float *tail_stack = nullptr;     // per-thread stacks of the tail launch
    size_t tail_floats = 0;
    hipEvent_t done[2] = {nullptr, nullptr};   // completion of the two replays in flight
    BatchPlan *d_plan = nullptr;     // injection plan of the current batch
    BatchPlan *h_plan = nullptr;     // pinned staging copy
    // the captured sweep group (graph[0]; graph[1] is unused) and, for
    // kernel timing, one set of event pairs per replay slot
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    EvArgs graph_args{};
    EventConfig graph_cfg{};
    static constexpr int MAX_TIMED = 64;
    hipEvent_t tev[2][2 * MAX_TIMED] = {};
    bool timing = false;             // measure the photon kernel with HIP events
    double photon_ms = 0.0;          // accumulated over the runs since the last reset
    uint64_t photon_launches = 0;
    SweepFork fork;                  // the electron side's stream and its fork/join events

    // Banks for a run of n_hist histories.  Without an explicit capacity a
    // bank holds twice what can be in flight at an injection (what is left
    // below the threshold plus the injection, neither more than the run's
    // histories), and at least 4 Mi particles.
    static size_t auto_capacity(uint64_t n_hist, const EventConfig &cfg) {
        const uint64_t inj = cfg.inject < n_hist ? cfg.inject : n_hist;
        const uint64_t thr = cfg.threshold < n_hist ? cfg.threshold : n_hist;
        const uint64_t want = 2 * (inj + thr);
        const uint64_t floor = (uint64_t)1 << 22;
        const uint64_t cap = want > floor ? want : floor;
        return (size_t)(cap < 0xFFFFF000ull ? cap : 0xFFFFF000ull);
    }
    hipError_t reserve(uint64_t n_hist, uint32_t, const EventConfig &cfg) {
        // a whole number of append segments per bank
        const size_t want = (cfg.capacity ? cfg.capacity : auto_capacity(n_hist, cfg)) / QSEG * QSEG;
        if (want > 0xFFFFFFF0ull || want < (size_t)QSEG) return hipErrorInvalidValue;
        hipError_t e = hipSuccess;
        if (!st) {
            e = hipMalloc(&st, sizeof(EvState));
            if (e != hipSuccess) return e;
            e = hipHostMalloc(&h_flags, 2 * FLAG_WORDS * sizeof(uint32_t));
            if (e != hipSuccess) return e;
            e = hipMalloc(&d_plan, sizeof(BatchPlan));
            if (e != hipSuccess) return e;
            e = hipHostMalloc(&h_plan, sizeof(BatchPlan));
            if (e != hipSuccess) return e;
            for (int k = 0; k < 2; ++k) {
                e = hipEventCreateWithFlags(&done[k], hipEventDisableTiming);
                if (e != hipSuccess) return e;
            }
            e = hipStreamCreateWithFlags(&fork.side, hipStreamNonBlocking);
            if (e != hipSuccess) return e;
            e = hipEventCreateWithFlags(&fork.fork, hipEventDisableTiming);
            if (e != hipSuccess) return e;
            e = hipEventCreateWithFlags(&fork.join, hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
        if (cap < want) {
            if (bank) (void)hipFree(bank);
            bank = nullptr;
            cap = 0;
            e = hipMalloc(&bank, want * N_SLOTS * EVW * sizeof(float));
            if (e != hipSuccess) return e;
            cap = (uint32_t)want;
            drop_graph();
        }
        return hipSuccess;
    }
    void drop_graph() {
        for (int k = 0; k < 2; ++k) {
            if (graph[k]) (void)hipGraphExecDestroy(graph[k]);
            graph[k] = nullptr;
        }
    }
    hipError_t make_timing_events() {
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < 2 * MAX_TIMED; ++i)
                if (!tev[k][i]) {
                    const hipError_t e = hipEventCreate(&tev[k][i]);
                    if (e != hipSuccess) return e;
                }
        return hipSuccess;
    }
    // photon kernel launches of the replay in `slot`, timed by the event
    // pairs its graph records on the stream
    hipError_t add_timing(int slot, uint32_t sweeps) {
        for (uint32_t k = 0; k < sweeps; ++k) {
            float ms = 0.0f;
            const hipError_t e = hipEventElapsedTime(&ms, tev[slot][2 * k], tev[slot][2 * k + 1]);
            if (e != hipSuccess) return e;
            photon_ms += ms;
            photon_launches += 1;
        }
        return hipSuccess;
    }
    void release() {
        drop_graph();
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < 2 * MAX_TIMED; ++i)
                if (tev[k][i]) (void)hipEventDestroy(tev[k][i]);
        if (bank) (void)hipFree(bank);
        if (st) (void)hipFree(st);
        if (h_flags) (void)hipHostFree(h_flags);
        if (tail_stack) (void)hipFree(tail_stack);
        tail_stack = nullptr;
        tail_floats = 0;
        if (d_plan) (void)hipFree(d_plan);
        for (int k = 0; k < 2; ++k)
            if (done[k]) (void)hipEventDestroy(done[k]);
        done[0] = done[1] = nullptr;
        if (fork.side) (void)hipStreamDestroy(fork.side);
        if (fork.fork) (void)hipEventDestroy(fork.fork);
        if (fork.join) (void)hipEventDestroy(fork.join);
        fork = SweepFork();
        if (h_plan) (void)hipHostFree(h_plan);
        bank = nullptr;
        st = nullptr;
        h_flags = nullptr;
        d_plan = nullptr;
        h_plan = nullptr;
        cap = 0;
    }
};

// blocks that keep every CU busy once, for the persistent flight kernels
template <class K>
inline uint32_t resident_grid(K kernel, uint32_t block, size_t lds) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (uint32_t)(per_cu * cus);
}

inline bool tab_in_lds(const Env &env) { return env.nclass <= LDS_MAX_CLASSES; }

inline size_t flight_lds(const Env &env) {
    return env_lds_bytes(env.nx, env.ny, env.nz, env.nclass, env.nmed, env.njob, tab_in_lds(env));
}

template <bool TAB, bool ADJ, bool GEN>
inline void launch_sweep_t(const EvArgs &a, const EventConfig &cfg, hipStream_t stream, hipEvent_t t0,
                           hipEvent_t t1, const SweepFork &f) {
    const dim3 g(cfg.grid), b(cfg.block), fb(cfg.flight_block);
    const size_t lds = flight_lds(a.env);
    hipLaunchKernelGGL(inject_fill, g, b, 0, stream, a);
    hipStream_t es = stream;
    if (f.side) {
        (void)hipEventRecord(f.fork, stream);
        (void)hipStreamWaitEvent(f.side, f.fork, 0);
        es = f.side;
    }
    if (t0) (void)hipEventRecord(t0, stream);
    hipLaunchKernelGGL((photon_kernel<TAB, ADJ, GEN>), dim3(cfg.photon_grid), fb, lds, stream, a, cfg.photon_chunk,
                       cfg.refill_min, cfg.refill_slots);
    if (t1) (void)hipEventRecord(t1, stream);
    hipLaunchKernelGGL((interaction_kernel<P_COMPTON, ADJ>), g, b, 0, stream, a);
    hipLaunchKernelGGL((electron_kernel<TAB, ADJ, GEN>), dim3(cfg.electron_grid), fb, lds, es, a, cfg.electron_chunk,
                       cfg.refill_min, cfg.refill_slots);
    hipLaunchKernelGGL((interaction_kernel<E_MOLLER, ADJ>), g, b, 0, es, a);
    if (f.side) {
        (void)hipEventRecord(f.join, f.side);
        (void)hipStreamWaitEvent(stream, f.join, 0);
    }
    hipLaunchKernelGGL(sweep_end, dim3(1), dim3(1), 0, stream, a.st, cfg.threshold, cfg.inject);
}

template <bool GEN>
inline void launch_sweep_f(const EvArgs &a, const EventConfig &cfg, hipStream_t stream, hipEvent_t t0,
                           hipEvent_t t1, const SweepFork &f) {
    const bool adj = a.adj != nullptr;
    if (tab_in_lds(a.env)) {
        if (adj) launch_sweep_t<true, true, GEN>(a, cfg, stream, t0, t1, f);
        else launch_sweep_t<true, false, GEN>(a, cfg, stream, t0, t1, f);
    } else {
        if (adj) launch_sweep_t<false, true, GEN>(a, cfg, stream, t0, t1, f);
        else launch_sweep_t<false, false, GEN>(a, cfg, stream, t0, t1, f);
    }
}

inline void launch_sweep(const EvArgs &a, const EventConfig &cfg, hipStream_t stream, hipEvent_t t0 = nullptr,
                         hipEvent_t t1 = nullptr, const SweepFork &f = SweepFork()) {
    if (a.env.fixes & ENV_GENERAL) launch_sweep_f<true>(a, cfg, stream, t0, t1, f);
    else launch_sweep_f<false>(a, cfg, stream, t0, t1, f);
}

constexpr uint32_t TAIL_BLOCK = 128;
constexpr int TAIL_DEPTH = 40;       // particles per thread's stack (the plain kernel's default)

// threads of the tail launch: one per particle up to `tail`, whole blocks
inline uint32_t tail_threads(const EventConfig &cfg) {
    return (cfg.tail + TAIL_BLOCK - 1) / TAIL_BLOCK * TAIL_BLOCK;
}

// The tail launch and the bookkeeping after it; `stack` holds
// tail_threads(cfg) x TAIL_DEPTH particles.
template <bool GEN>
inline void launch_tail_g(const EvArgs &a, const EventConfig &cfg, float *stack, hipStream_t stream) {
    const dim3 g(tail_threads(cfg) / TAIL_BLOCK), b(TAIL_BLOCK);
    const size_t lds = flight_lds(a.env);
    const bool adj = a.adj != nullptr;
    if (tab_in_lds(a.env)) {
        if (adj) hipLaunchKernelGGL((tail_kernel<true, true, GEN>), g, b, lds, stream, a, stack, TAIL_DEPTH);
        else hipLaunchKernelGGL((tail_kernel<true, false, GEN>), g, b, lds, stream, a, stack, TAIL_DEPTH);
    } else {
        if (adj) hipLaunchKernelGGL((tail_kernel<false, true, GEN>), g, b, lds, stream, a, stack, TAIL_DEPTH);
        else hipLaunchKernelGGL((tail_kernel<false, false, GEN>), g, b, lds, stream, a, stack, TAIL_DEPTH);
    }
    hipLaunchKernelGGL(tail_end, dim3(1), dim3(1), 0, stream, a.st);
}

inline void launch_tail(const EvArgs &a, const EventConfig &cfg, float *stack, hipStream_t stream) {
    if (a.env.fixes & ENV_GENERAL) launch_tail_g<true>(a, cfg, stack, stream);
    else launch_tail_g<false>(a, cfg, stack, stream);
}

template <bool TAB, bool ADJ, bool GEN>
inline void flight_grids_t(EventConfig &cfg, size_t lds) {
    if (!cfg.photon_grid) cfg.photon_grid = resident_grid(photon_kernel<TAB, ADJ, GEN>, cfg.flight_block, lds);
    if (!cfg.electron_grid) cfg.electron_grid = resident_grid(electron_kernel<TAB, ADJ, GEN>, cfg.flight_block, lds);
}

template <bool GEN>
inline void flight_grids_f(EventConfig &cfg, const Env &env, bool adj) {
    const size_t lds = flight_lds(env);
    if (tab_in_lds(env)) {
        if (adj) flight_grids_t<true, true, GEN>(cfg, lds);
        else flight_grids_t<true, false, GEN>(cfg, lds);
    } else {
        if (adj) flight_grids_t<false, true, GEN>(cfg, lds);
        else flight_grids_t<false, false, GEN>(cfg, lds);
    }
}

// grid sizes of the persistent flight kernels that fill the device once
inline void flight_grids(EventConfig &cfg, const Env &env, bool adj) {
    if (env.fixes & ENV_GENERAL) flight_grids_f<true>(cfg, env, adj);
    else flight_grids_f<false>(cfg, env, adj);
}

inline bool same_args(const EvArgs &x, const EvArgs &y) {
    return x.bank == y.bank && x.cap == y.cap && x.st == y.st && x.src == y.src && x.plan == y.plan &&
           x.job_plan == y.job_plan && x.nvox == y.nvox && x.env.fields == y.env.fields && x.env.njob == y.env.njob &&
           x.env.nclass == y.env.nclass && x.env.nmed == y.env.nmed &&
           x.dose == y.dose && x.hist_edep == y.hist_edep && x.hist_steps == y.hist_steps &&
           x.beamlet == y.beamlet && x.bweight == y.bweight && x.adj == y.adj && x.grad == y.grad &&
           x.counters == y.counters && x.env.mt == y.env.mt && x.env.rcls == y.env.rcls && x.env.rtab == y.env.rtab && x.env.seed == y.env.seed &&
           x.env.xb == y.env.xb && x.env.yb == y.env.yb && x.env.zb == y.env.zb && x.env.nx == y.env.nx &&
           x.env.ny == y.env.ny && x.env.nz == y.env.nz && x.env.pge == y.env.pge && x.env.pgmfp == y.env.pgmfp &&
           x.env.pcohe == y.env.pcohe && x.env.mlc == y.env.mlc && x.env.endnodes == y.env.endnodes &&
           x.env.mlc_next == y.env.mlc_next && x.env.jaws == y.env.jaws && x.env.jawx_next == y.env.jawx_next &&
           x.env.bw == y.env.bw && x.env.block_next == y.env.block_next && x.env.wedge_nver == y.env.wedge_nver &&
           x.env.fixes == y.env.fixes &&
           x.tally_batches == y.tally_batches;
}

inline bool same_cfg(const EventConfig &x, const EventConfig &y) {
    return x.block == y.block && x.flight_block == y.flight_block && x.grid == y.grid && x.photon_grid == y.photon_grid &&
           x.electron_grid == y.electron_grid && x.photon_chunk == y.photon_chunk &&
           x.electron_chunk == y.electron_chunk && x.refill_min == y.refill_min && x.inject == y.inject &&
           x.threshold == y.threshold && x.sweeps_per_graph == y.sweeps_per_graph && x.refill_slots == y.refill_slots &&
           x.sweep_streams == y.sweep_streams && x.tail == y.tail;
}

// Transport the histories of one batch (bk.h_plan, filled by the caller)
// through the banks.  The sweep group is captured once per argument set and
// replayed; the host polls one flag per replay.
// beamlet weights and adjoint buffers of a run (all optional)
struct BeamletArgs {
    const uint32_t *beamlet = nullptr;
    const float *bweight = nullptr;
    const double *adj = nullptr;
    double *grad = nullptr;
};

inline hipError_t run_event_batch(EventBanks &bk, const Env &env, const gegs_source *src, const uint32_t *job_plan,
                                  size_t nvox, double *dose, double *hist_edep, uint32_t *hist_steps,
                                  unsigned long long *counters, const BeamletArgs &bl, const EventConfig &cfg_in,
                                  hipStream_t stream, uint32_t tally_batches = 1, uint32_t max_sweeps = 1000000) {
    EventConfig cfg = cfg_in;
    flight_grids(cfg, env, bl.adj != nullptr);
    // the work counters `taken` run past the bank size by at most one
    // reservation per wavefront; keep them inside 32 bits
    const uint64_t waves = (uint64_t)(cfg.photon_grid > cfg.electron_grid ? cfg.photon_grid : cfg.electron_grid) *
                           (cfg.flight_block / 64);
    if ((uint64_t)bk.cap + waves * cfg.refill_slots >= 0xFFFFFFFFull) return hipErrorInvalidValue;
    EvArgs a;
    a.env = env;
    a.plan = bk.d_plan;
    a.job_plan = job_plan;
    a.nvox = nvox;
    a.bank = bk.bank;
    a.cap = bk.cap;
    a.seg_cap = bk.cap / QSEG;
    a.st = bk.st;
    a.src = src;
    a.dose = dose;
    a.hist_edep = hist_edep;
    a.hist_steps = hist_steps;
    a.counters = counters;
    a.beamlet = bl.beamlet;
    a.bweight = bl.bweight;
    a.adj = bl.adj;
    a.grad = bl.grad;
    a.tally_batches = tally_batches;
    hipError_t e;
    // With kernel timing the sweeps are launched directly on the stream,
    // with an event pair around every photon kernel launch (HIP reports no
    // elapsed time for events recorded by graph nodes); otherwise the sweep
    // group is captured once per argument set and replayed.
    const bool timed = bk.timing && cfg.sweeps_per_graph <= (uint32_t)EventBanks::MAX_TIMED;
    const SweepFork fork = cfg.sweep_streams == 2 ? bk.fork : SweepFork();
    if (timed) {
        e = bk.make_timing_events();
        if (e != hipSuccess) return e;
    } else if (!bk.graph[0] || !same_args(a, bk.graph_args) || !same_cfg(cfg, bk.graph_cfg)) {
        bk.drop_graph();
        hipGraph_t g;
        e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        for (uint32_t k = 0; k < cfg.sweeps_per_graph; ++k) launch_sweep(a, cfg, stream, nullptr, nullptr, fork);
        e = hipStreamEndCapture(stream, &g);
        if (e != hipSuccess) return e;
        e = hipGraphInstantiate(&bk.graph[0], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return e;
        bk.graph_args = a;
        bk.graph_cfg = cfg;
    }
    e = hipMemcpyAsync(bk.d_plan, bk.h_plan, sizeof(BatchPlan), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_begin, dim3(1), dim3(1), 0, stream, bk.st, 0u, bk.h_plan->total, cfg.threshold, cfg.inject);
    // Two replays in flight: while the host waits for the flag of replay k,
    // replay k+1 is already queued, so the GPU does not idle during the host
    // round trip.  The replay queued after the last busy one finds every bank
    // empty and no record left; its kernels return at once.
    uint32_t sweeps = 0;
    auto launch = [&](int slot) -> hipError_t {
        hipError_t r = hipSuccess;
        if (timed) {
            for (uint32_t k = 0; k < cfg.sweeps_per_graph; ++k)
                launch_sweep(a, cfg, stream, bk.tev[slot][2 * k], bk.tev[slot][2 * k + 1], fork);
            r = hipGetLastError();
        } else {
            r = hipGraphLaunch(bk.graph[0], stream);
        }
        if (r != hipSuccess) return r;
        r = hipMemcpyAsync(bk.h_flags + FLAG_WORDS * slot, &bk.st->active, FLAG_WORDS * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, stream);
        if (r != hipSuccess) return r;
        return hipEventRecord(bk.done[slot], stream);
    };
    // The tail launch finishes the run once every record is injected and at
    // most cfg.tail particles are in flight (not in a timed pass, whose
    // photon steps must all be photon-kernel steps).
    const bool use_tail = cfg.tail && !timed;
    if (use_tail) {
        const size_t need = (size_t)tail_threads(cfg) * TAIL_DEPTH * PART_WORDS;
        if (bk.tail_floats < need) {
            if (bk.tail_stack) (void)hipFree(bk.tail_stack);
            bk.tail_stack = nullptr;
            bk.tail_floats = 0;
            e = hipMalloc(&bk.tail_stack, need * sizeof(float));
            if (e != hipSuccess) return e;
            bk.tail_floats = need;
        }
    }
    auto tail_ready = [&](int slot) {
        const uint32_t *f = bk.h_flags + FLAG_WORDS * slot;
        // every record injected (src_next >= src_last, no injection planned), inflight
        return use_tail && f[3] >= f[4] && f[6] == 0 && f[7] <= cfg.tail;
    };
    if ((e = launch(0)) != hipSuccess) return e;
    if ((e = launch(1)) != hipSuccess) return e;
    for (int k = 0;; ++k) {
        const int slot = k & 1, other = slot ^ 1;
        e = hipEventSynchronize(bk.done[slot]);
        if (e != hipSuccess) return e;
        if (timed && (e = bk.add_timing(slot, cfg.sweeps_per_graph)) != hipSuccess) return e;
        sweeps += cfg.sweeps_per_graph;
        if (!bk.h_flags[FLAG_WORDS * slot] || tail_ready(slot)) {
            // the replay queued behind this one finds nothing (or a tail's
            // worth) left to do; wait for it so that its launches are counted
            e = hipEventSynchronize(bk.done[other]);
            if (e != hipSuccess) return e;
            if (timed && (e = bk.add_timing(other, cfg.sweeps_per_graph)) != hipSuccess) return e;
            if (bk.h_flags[FLAG_WORDS * other]) {
                launch_tail(a, cfg, bk.tail_stack, stream);
                e = hipGetLastError();
                if (e != hipSuccess) return e;
                e = hipStreamSynchronize(stream);
                if (e != hipSuccess) return e;
            }
            break;
        }
        if (sweeps >= max_sweeps) {
            (void)hipStreamSynchronize(stream);
            return hipErrorLaunchTimeOut;
        }
        if ((e = launch(slot)) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace gegs
