// plain.hpp -- the plain per-history kernel (the baseline method).
//
// One thread transports one history at a time with its own LIFO particle
// stack, exactly the order the oracle uses; the stack lives in HBM in a
// slot-major, thread-minor layout so that the lanes of a wavefront touching
// the same slot depth hit consecutive words.  Histories are handed out by an
// atomic counter, so threads that finish early pick up more work.
#pragma once
#include "transport.hpp"

namespace gegs {

constexpr int PART_WORDS = 16;

struct StackSink {
    float *base;          // [slot][word][thread]
    uint32_t stride;      // threads in the grid
    uint32_t tid;
    int sp;
    int depth;
    uint32_t overflow;

    __device__ __forceinline__ float *at(int slot, int word) const {
        return base + ((size_t)slot * PART_WORDS + word) * stride + tid;
    }
    __device__ __forceinline__ void push(const Part &p) {
        if (sp >= depth) {
            overflow++;
            return;
        }
        const int k = sp++;
        *at(k, 0) = p.e;
        *at(k, 1) = p.x;
        *at(k, 2) = p.y;
        *at(k, 3) = p.z;
        *at(k, 4) = p.u;
        *at(k, 5) = p.v;
        *at(k, 6) = p.w;
        *at(k, 7) = p.wt;
        *at(k, 8) = p.dnear;
        *at(k, 9) = __uint_as_float(p.region);
        *at(k, 10) = __uint_as_float((uint32_t)(p.charge + 1) | ((uint32_t)p.status << 8) | ((uint32_t)p.module << 16) |
                                     ((uint32_t)p.spos << 24));
        *at(k, 11) = __uint_as_float((uint32_t)p.lin);
        *at(k, 12) = __uint_as_float((uint32_t)(p.lin >> 32));
        *at(k, 13) = __uint_as_float(p.blk);
        *at(k, 14) = __uint_as_float(p.job);
        *at(k, 15) = __uint_as_float((uint32_t)(p.leaf + 128) | ((uint32_t)(p.bank + 1) << 8));
    }
    __device__ __forceinline__ Part pop() {
        const int k = --sp;
        Part p;
        p.e = *at(k, 0);
        p.x = *at(k, 1);
        p.y = *at(k, 2);
        p.z = *at(k, 3);
        p.u = *at(k, 4);
        p.v = *at(k, 5);
        p.w = *at(k, 6);
        p.wt = *at(k, 7);
        p.dnear = *at(k, 8);
        p.region = __float_as_uint(*at(k, 9));
        const uint32_t pk = __float_as_uint(*at(k, 10));
        p.charge = (int32_t)(pk & 0xffu) - 1;
        p.status = (int32_t)((pk >> 8) & 0xffu);
        p.module = (int32_t)((pk >> 16) & 0xffu);
        p.spos = (int32_t)(pk >> 24);
        p.lin = (uint64_t)__float_as_uint(*at(k, 11)) | ((uint64_t)__float_as_uint(*at(k, 12)) << 32);
        p.blk = __float_as_uint(*at(k, 13));
        p.job = __float_as_uint(*at(k, 14));
        const uint32_t lb = __float_as_uint(*at(k, 15));
        p.leaf = (int32_t)(lb & 0xffu) - 128;
        p.bank = (int32_t)(lb >> 8) - 1;
        return p;
    }
};

struct PlainArgs {
    Env env;
    const gegs_source *src;
    uint32_t first, last;        // source records [first, last)
    uint64_t hist_offset;        // global index of record 0
    uint32_t *next;              // work counter of this launch, starts at 0
    float *stack;
    int depth;
    double *dose;
    double *hist_edep;           // optional, indexed by record
    uint32_t *hist_steps;        // optional
    unsigned long long *counters;   // gegs_counters layout
};

__device__ __forceinline__ Part source_particle(const gegs_source &r, int module) {
    Part p;
    p.e = r.e; p.x = r.x; p.y = r.y; p.z = r.z;
    p.u = r.u; p.v = r.v; p.w = r.w; p.wt = r.wt;
    p.charge = r.charge;
    p.region = 0;
    p.module = module;
    p.dnear = INFI_DIST;
    p.blk = 0;
    p.job = 0;
    p.leaf = OUTMLC;
    p.bank = -1;
    p.spos = 4;
    p.status = r.charge == 0 ? P_PHOTON_STEP : E_ELECTRON_STEP;
    return p;
}

template <bool TAB, bool GEN>
__global__ void __launch_bounds__(128) GEGS_WAVES_PER_EU(2) plain_kernel(PlainArgs a) {
    extern __shared__ float lds_env[];
    Env env = a.env;
    env.fixes = kernel_fixes<GEN>(env.fixes);
    stage_env<TAB>(env, lds_env);
    StackSink sink;
    sink.base = a.stack;
    sink.stride = gridDim.x * blockDim.x;
    sink.tid = blockIdx.x * blockDim.x + threadIdx.x;
    sink.depth = a.depth;
    sink.overflow = 0;
    Tally t;
    t.dose = a.dose;
    t.photon_steps = t.electron_steps = t.gathers = t.deposits = 0;
    uint32_t pops[N_EVENTS];
#pragma unroll
    for (int k = 0; k < N_EVENTS; ++k) pops[k] = 0;
    uint32_t nhist = 0;
    for (;;) {
        const uint32_t k = atomicAdd(a.next, 1u);
        if (k >= a.last - a.first) break;
        const uint32_t i = a.first + k;
        Part p = source_particle(a.src[i], env.fields[0].module_first);
        p.lin = primary_lineage(a.hist_offset + i);
        t.hist_edep = 0.0;
        t.steps = 0;
        sink.sp = 0;
        sink.push(p);
        while (sink.sp > 0) {
            Part q = sink.pop();
            Stream s;
            s.begin(q, env.seed);
#pragma unroll
            for (int k = 0; k < N_EVENTS; ++k) pops[k] += (q.status == k);
            handle(env, s, q, sink, t);
        }
        if (a.hist_edep) a.hist_edep[i] = t.hist_edep;
        if (a.hist_steps) a.hist_steps[i] = t.steps;
        nhist++;
    }
    // one atomic per wavefront and counter
    unsigned long long *c = a.counters;
    uint32_t vals[16];
    vals[0] = t.photon_steps;
    vals[1] = t.electron_steps;
    vals[2] = t.gathers;
    vals[3] = t.deposits;
#pragma unroll
    for (int k = 0; k < N_EVENTS; ++k) vals[4 + k] = pops[k];
    vals[14] = nhist;
    vals[15] = sink.overflow;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        uint32_t v = vals[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (__lane_id() == 0 && v) atomicAdd(c + k, (unsigned long long)v);
    }
}

// sampler check: one handler invocation per input particle, pushes listed
struct ListSink {
    gegs_particle *out;
    uint32_t max_out;
    uint32_t n;
    __device__ __forceinline__ void push(const Part &p) {
        if (n < max_out) {
            gegs_particle &o = out[n];
            o.e = p.e; o.x = p.x; o.y = p.y; o.z = p.z; o.u = p.u; o.v = p.v; o.w = p.w;
            o.wt = p.wt; o.dnear = p.dnear; o.region = p.region; o.charge = p.charge;
            o.status = p.status; o.lin = p.lin; o.blk = p.blk; o.pad = 0;
        }
        n++;
    }
};

__global__ void sample_kernel(Env env, const gegs_particle *in, uint32_t n, gegs_particle *out, uint32_t max_out,
                              uint32_t *n_out, double *edep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const gegs_particle r = in[i];
    Part p;
    p.e = r.e; p.x = r.x; p.y = r.y; p.z = r.z; p.u = r.u; p.v = r.v; p.w = r.w; p.wt = r.wt;
    p.dnear = r.dnear; p.region = r.region; p.charge = r.charge; p.status = r.status;
    p.module = env.fields[0].module_first; p.lin = r.lin; p.blk = r.blk; p.spos = 4; p.job = 0;
    p.leaf = OUTMLC; p.bank = -1;
    ListSink sink{out + (size_t)i * max_out, max_out, 0};
    Tally t;
    t.dose = nullptr;
    t.hist_edep = 0.0;
    t.steps = 0;
    t.photon_steps = t.electron_steps = t.gathers = t.deposits = 0;
    Stream s;
    s.begin(p, env.seed);
    handle(env, s, p, sink, t);
    n_out[i] = sink.n;
    edep[i] = t.hist_edep;
}

// batch boundary: fold the batch tally into the run totals and clear it
__global__ void fold_batch(double *chunk, double *dose, double *dose2, size_t n) {
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x) {
        const double d = chunk[v];
        if (d != 0.0) {
            dose[v] += d;
            dose2[v] += d * d;
            chunk[v] = 0.0;
        }
    }
}

// dose and dose2 of n_rows x nvox from nb per-batch rows each ([row][batch][voxel]),
// clearing the batch rows
__global__ void fold_batches(double *chunk, double *dose, double *dose2, size_t nvox, size_t n_rows, uint32_t nb) {
    const size_t n = nvox * n_rows;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x) {
        const size_t row = v / nvox, vox = v - row * nvox;
        double s = 0.0, s2 = 0.0;
        for (uint32_t b = 0; b < nb; ++b) {
            double *c = chunk + (row * nb + b) * nvox + vox;
            const double d = *c;
            if (d != 0.0) {
                s += d;
                s2 += d * d;
                *c = 0.0;
            }
        }
        dose[v] += s;
        dose2[v] += s2;
    }
}

__global__ void fill_zero_u32(uint32_t *p, size_t n) {
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
        p[v] = 0u;
}

__global__ void fill_zero_u64(unsigned long long *p, size_t n) {
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
        p[v] = 0ull;
}

__global__ void fill_zero(double *p, size_t n) {
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
        p[v] = 0.0;
}

// FP64 atomic adds at the given voxel indices, `reps` passes over them: the
// same hardware atomic the dose tally uses, for a measured peak rate
__global__ void atomic_probe(double *acc, const uint32_t *idx, uint64_t n, uint32_t reps) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t r = 0; r < reps; ++r)
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            unsafeAtomicAdd(acc + idx[i], 1.0);
}

}  // namespace gegs
