// pp_lib.hip - libphotonprop_hip.so: gfx950 photon propagation kernels and the
// C ABI declared in include/photonprop_hip.h.
//
// Kernel design (one photon per lane, photons grid-strided over the launch):
//   * targets (float4 x,y,z,r^2) and the spectrum's inverse-CDF knots are
//     staged in LDS once per workgroup; target reads are wave-uniform
//     broadcasts;
//   * every photon draws from its own Philox4x32-10 stream (counter =
//     (word, photon id, source)), so hits do not depend on the launch shape;
//   * one Philox block per step supplies all four uniforms of a step;
//   * hits are compacted per wave: ballot + mbcnt rank, one 64-bit atomic per
//     wave, coalesced SoA stores; slots past capacity are dropped and flagged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pp_device.h"
#include "pp_grid_build.h"
#include "pp_hits.h"
#include "pp_search.h"

#include <rocprim/device/device_radix_sort.hpp>

using pp::f3;

// ---------------------------------------------------------------------------
// device data
// ---------------------------------------------------------------------------
namespace {

using pp::HitRecord;
using pp::kMaxPhotonsPerSource;
using pp::kPidBits;
constexpr int kMaxTargets = 65535;

struct HitCols {
    HitRecord* rec;
    uint64_t capacity;
};

struct Counters {
    unsigned long long n_hits;    // reserved slots (may exceed capacity)
    unsigned long long n_ph_sim;  // completed photons
    unsigned int err;             // 1 = overflow
    unsigned int pad;
    // work counts of the instrumented kernel (pp_config.count_work)
    unsigned long long steps, tests, cells, walks;
    unsigned long long next_chunk;  // work-stealing counter of the batch kernel
};

struct KParams {
    const pp_source* sources;  // device copy of the source list (one launch per source)
    pp::Medium med;
    uint32_t k0, k1, src_ix;
    int32_t n_steps;
    uint64_t n_photons;
    const float4* targets;
    int32_t n_targets;
    // uniform grid over the targets (GridSearch)
    pp::GridDesc grid;
    const float4* cells;
    const uint16_t* cell_items;
    int32_t spec_n;
    int32_t spec_map;
    float spec_mono;
    float spec_nf, spec_lastf;  // float(spec_n), float(spec_n - 1)
    const float* spec;
    HitCols hits;
    Counters* counters;
    uint32_t queue_off;  // LDS byte offset of the waves' walk queues (grid searches)
};

// ---------------------------------------------------------------------------
// photon emission   src/photon_prop_cuda.jl:158-276
// ---------------------------------------------------------------------------
struct PhotonInit {
    f3 pos, dir;
    float time, wl;
    pp::MediumValues mv;
};

// p + s d, each component one FMA (the same bits wherever it is inlined)
__device__ __forceinline__ f3 add_scaled(f3 p, float s, f3 d) {
    return pp::make3(fmaf(s, d.x, p.x), fmaf(s, d.y, p.y), fmaf(s, d.z, p.z));
}

// sample_cherenkov_direction[_no_smear|_custom_smear]  :80-111, :134-154
__device__ __forceinline__ f3 cherenkov_dir(const pp_source& src, const pp::MediumValues& mv,
                                            const pp::Stream& s, int smear) {
    const uint4 x = s.block(PP_CTR_INIT_EXTRA | 1u);
    f3 ph = pp::dir_from_cos_phi(mv.cos_c, mv.sin_c, pp::u01(x.x));
    if (smear == 1) {
        ph = pp::rot_from_ez(pp::cherenkov_track_dir(pp::u01(x.y), pp::u01(x.z)), ph);
    } else if (smear == 2) {
        const float gx = pp::gamma_sample(src.beta_a, 1.0f, s, PP_SUB_BETA_X);
        const float gy = pp::gamma_sample(src.beta_b, 1.0f, s, PP_SUB_BETA_Y);
        const float tc = (gx / (gx + gy)) * 2.0f - 1.0f;
        const float ts = pp::sqrt_hw((1.0f - tc) * (1.0f + tc));
        ph = pp::rot_from_ez(pp::dir_from_cos_phi(tc, ts, pp::u01(x.y)), ph);
    }
    const f3 axis = pp::make3(src.direction[0], src.direction[1], src.direction[2]);
    return pp::rot_from_ez(axis, ph);
}

__device__ __forceinline__ PhotonInit init_photon(const KParams& p, const pp_source& src,
                                                  const float* spec, const pp::Stream& s) {
    const uint4 r = s.block(PP_CTR_INIT);
    PhotonInit st;
    st.wl = pp::spectrum_sample(spec, p.spec_n, p.spec_nf, p.spec_lastf, p.spec_map, p.spec_mono,
                                pp::u01(r.x));
    st.mv = pp::medium_values(p.med, st.wl);
    st.pos = pp::make3(src.position[0], src.position[1], src.position[2]);
    st.time = src.time;
    const f3 axis = pp::make3(src.direction[0], src.direction[1], src.direction[2]);
    switch (src.kind) {
    case PP_SRC_POINTLIKE_ISOTROPIC:
        st.dir = pp::isotropic_dir(pp::u01(r.y), pp::u01(r.z));
        break;
    case PP_SRC_POINTLIKE_TIME_RANGE:
        st.dir = pp::isotropic_dir(pp::u01(r.y), pp::u01(r.z));
        st.time = src.time + pp::u01(r.w) * (src.time_end - src.time);
        break;
    case PP_SRC_AXICONE: {
        float theta = src.angle;
        if (src.beam_divergence > 0.0f)
            theta += pp::normal_sample(s, PP_SUB_NORMAL) * src.beam_divergence;
        float sn, cs;
        sincosf(theta, &sn, &cs);
        st.dir = pp::rot_from_ez(axis, pp::dir_from_cos_phi(cs, sn, pp::u01(r.y)));
        break;
    }
    case PP_SRC_COLLIMATED_ISOTROPIC: {
        const float ct = src.cut_angle_cos + pp::u01(r.z) * (1.0f - src.cut_angle_cos);
        st.dir = pp::dir_from_cos_phi(ct, pp::sqrt_hw((1.0f - ct) * (1.0f + ct)), pp::u01(r.y));
        break;
    }
    case PP_SRC_PENCIL:
        if (src.beam_divergence > 0.0f) {
            const float theta =
                pp::sqrt_hw(pp::u01(r.y) * (src.beam_divergence * src.beam_divergence));
            float sn, cs;
            sincosf(theta, &sn, &cs);
            st.dir = pp::rot_from_ez(axis, pp::dir_from_cos_phi(cs, sn, pp::u01(r.z)));
        } else {
            st.dir = axis;
        }
        break;
    case PP_SRC_EXTENDED_CHERENKOV:
    case PP_SRC_EXTENDED_CHERENKOV_NO_SMEAR:
    case PP_SRC_EXTENDED_CHERENKOV_CUSTOM_SMEAR: {
        const float lp =
            pp::gamma_sample(src.long_a, 1.0f / src.long_b, s, PP_SUB_LONG_GAMMA) * src.long_lrad;
        st.pos = add_scaled(st.pos, lp, axis);
        st.time = src.time + lp * (1.0f / pp::kCVacMNs);
        const int smear = src.kind == PP_SRC_EXTENDED_CHERENKOV ? 1
                          : src.kind == PP_SRC_EXTENDED_CHERENKOV_NO_SMEAR ? 0 : 2;
        st.dir = cherenkov_dir(src, st.mv, s, smear);
        break;
    }
    case PP_SRC_POINTLIKE_CHERENKOV:
        st.dir = cherenkov_dir(src, st.mv, s, 1);
        break;
    case PP_SRC_CHERENKOV_TRACK:
    default: {
        const float along = pp::u01(r.y) * src.length;
        st.pos = add_scaled(st.pos, along, axis);
        st.time = src.time + along * (1.0f / pp::kCVacMNs);
        st.dir = cherenkov_dir(src, st.mv, s, 1);
        break;
    }
    }
    return st;
}

// ---------------------------------------------------------------------------
// wave-compacted hit store
// ---------------------------------------------------------------------------
struct HitRec {
    f3 pos, dir, init_dir;
    float wl, time, dist;
    int32_t module_ix, n_steps;
    uint64_t pid;
};

// Returns false if this lane's record did not fit (overflow).
__device__ __forceinline__ bool store_hits(bool has_hit, const HitRec& h, uint32_t src_ix,
                                           const HitCols& out, Counters* ctr) {
    const uint64_t mask = __ballot(has_hit);
    if (mask == 0) return true;
    const uint32_t lane = __lane_id();
    const uint32_t rank =
        __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(&ctr->n_hits, (unsigned long long)__popcll(mask));
    const uint32_t blo = __builtin_amdgcn_readlane((uint32_t)base, leader);
    const uint32_t bhi = __builtin_amdgcn_readlane((uint32_t)(base >> 32), leader);
    base = ((unsigned long long)bhi << 32) | blo;
    if (!has_hit) return true;
    const uint64_t i = base + rank;
    if (i >= out.capacity) return false;
    float4* r = reinterpret_cast<float4*>(out.rec + i);
    r[0] = make_float4(h.pos.x, h.pos.y, h.pos.z, h.dir.x);
    r[1] = make_float4(h.dir.y, h.dir.z, h.init_dir.x, h.init_dir.y);
    r[2] = make_float4(h.init_dir.z, h.wl, h.time, h.dist);
    const uint64_t id = h.pid | ((uint64_t)(uint32_t)h.module_ix << kPidBits);
    r[3] = make_float4(__uint_as_float((uint32_t)id), __uint_as_float((uint32_t)(id >> 32)),
                       __int_as_float(h.n_steps), __uint_as_float(src_ix));
    return true;
}

// ---------------------------------------------------------------------------
// propagation kernel: cuda_propagate_photons!   src/photon_prop_cuda.jl:429-549
// ---------------------------------------------------------------------------
using pp::lds_bytes;

// kSearchGridGlobal: the grid search on tables read from global memory (L2
// and the vector cache) instead of LDS copies, for target lists whose records
// and grid tables exceed the LDS
enum { kSearchBrute = 0, kSearchGrid = 1, kSearchMixed = 2, kSearchGridGlobal = 3 };
// the grid searches walk in batches (transport_group); the every-target
// scans test each step at once (transport_photon)
template <int SEARCH>
constexpr bool kDeferredWalk = SEARCH == kSearchGrid || SEARCH == kSearchGridGlobal;

// LDS tables of one workgroup: targets, grid cells and items, spectrum knots
struct Lds {
    const float4* tgt;
    const float4* cells;
    const uint16_t* items;
    float* spec;
};

template <int SEARCH>
__device__ __forceinline__ Lds stage_lds(const KParams& p) {
    extern __shared__ float4 lds[];
    Lds l;
    if constexpr (SEARCH == kSearchGridGlobal) {
        // only the spectrum is staged; the search reads the device tables
        l.tgt = p.targets;
        l.cells = p.cells;
        l.items = p.cell_items;
        l.spec = reinterpret_cast<float*>(lds);
        for (int i = threadIdx.x; i < p.spec_n; i += blockDim.x) l.spec[i] = p.spec[i];
        __syncthreads();
        return l;
    }
    // a mixed (sphere and planar) list has two records per target
    const int n_rec = SEARCH == kSearchMixed ? 2 * p.n_targets : p.n_targets;
    float4* const tgt = lds;
    float4* const cells = lds + n_rec;
    uint16_t* const items =
        reinterpret_cast<uint16_t*>(cells + (SEARCH == kSearchGrid ? 2 * p.grid.n_cells : 0));
    l.tgt = tgt;
    l.cells = cells;
    l.items = items;
    l.spec = reinterpret_cast<float*>(items + (SEARCH == kSearchGrid ? ((p.grid.n_items + 1) & ~1) : 0));
    for (int i = threadIdx.x; i < n_rec; i += blockDim.x) tgt[i] = p.targets[i];
    if (SEARCH == kSearchGrid) {
        for (int i = threadIdx.x; i < 2 * p.grid.n_cells; i += blockDim.x) cells[i] = p.cells[i];
        for (int i = threadIdx.x; i < p.grid.n_items; i += blockDim.x) items[i] = p.cell_items[i];
    }
    for (int i = threadIdx.x; i < p.spec_n; i += blockDim.x) l.spec[i] = p.spec[i];
    __syncthreads();
    return l;
}

template <int SEARCH>
using SearchT = typename std::conditional<
    SEARCH == kSearchGrid || SEARCH == kSearchGridGlobal, pp::GridSearch,
    typename std::conditional<SEARCH == kSearchMixed, pp::MixedSearch, pp::BruteSearch>::type>::type;

template <int SEARCH>
__device__ __forceinline__ SearchT<SEARCH> make_search(const KParams& p, const Lds& l) {
    if constexpr (SEARCH == kSearchGrid || SEARCH == kSearchGridGlobal) {
        return pp::GridSearch{l.tgt, l.cells, l.items, p.grid};
    } else if constexpr (SEARCH == kSearchMixed) {
        return pp::MixedSearch{l.tgt, p.n_targets};
    } else {
        return pp::BruteSearch{l.tgt, p.n_targets};
    }
}

// position, distance and time after moving d along dir
__device__ __forceinline__ void move_by(f3& pos, float& dist, float& time, f3 dir, float d,
                                        float inv_c) {
    pos = add_scaled(pos, d, dir);
    dist += d;
    time = fmaf(d, inv_c, time);
}

// One photon's life: emission, then up to n_steps steps of
// src/photon_prop_cuda.jl:486-539. Returns true on a hit; h holds the record.
template <bool COUNT, typename Search>
__device__ __forceinline__ bool transport_photon(const KParams& p, const pp_source& src,
                                                 uint32_t src_ix, uint64_t pid, const float* spec,
                                                 const Search& search, const pp::Phase& phase,
                                                 pp::Work& work, unsigned long long& n_steps_done,
                                                 HitRec& h) {
    pp::Stream s{p.k0, p.k1, (uint32_t)pid, (uint32_t)(pid >> 32), src_ix};
    const PhotonInit st = init_photon(p, src, spec, s);
    const pp::StepStream ss(s);
    f3 pos = st.pos, dir = st.dir;
    float time = st.time, dist = 0.0f;
    const float ln2_sca = st.mv.sca_len * pp::kLn2, inv_c = st.mv.inv_c_grp;
    bool hit = false;
    int32_t module_ix = 0, n_at = 0;
    for (int32_t step = 1; step <= p.n_steps; ++step) {
        const uint4 r = ss.block((uint32_t)step);
        const float step_size = pp::step_length(pp::u01_open0(r.x), ln2_sca);
        // first target in index order whose near root lies inside the step
        float dt = 0.0f;
        const int found = search.template find<COUNT>(pos, dir, step_size, dt, work);
        if (COUNT) ++n_steps_done;
        if (found >= 0) {
            move_by(pos, dist, time, dir, dt, inv_c);
            hit = true;
            module_ix = found + 1;
            n_at = step;
            break;
        }
        move_by(pos, dist, time, dir, step_size, inv_c);
        const float ct = pp::phase_cos_words(phase, r.y, r.z);
        dir = pp::update_direction(dir, ct, pp::u01(r.w));
    }
    h = HitRec{pos, dir, st.dir, st.wl, time, dist, module_ix, n_at, pid};
    return hit;
}

// ---------------------------------------------------------------------------
// Deferred walk of the grid searches. A wave transports its 64 photons step
// by step as in transport_photon, but a step whose segment comes near the
// grid box (its end points' box meets it) does not walk the grid at once: the
// segment becomes a job in the wave's LDS queue, the photon moves on as if
// nothing were hit, and whenever 64 jobs are queued every lane of the wave
// walks one of them. The walks of a step run with 20-30% of the lanes in the
// one-photon schedule (the others are outside the grid or done); in a batch
// every lane walks. A job that hits records (step << 16 | target) in its
// photon's LDS slot with an atomic minimum; the earliest step wins, which is
// the step at which transport_photon stops (its later steps were never taken
// by the reference and their jobs only lose the minimum), so a photon need
// not stop at a hit it does not yet know of. A hit's record holds the photon
// id, module and step; replay_hits_kernel then emits the few photons with a
// hit again and replays their steps up to it (same random stream, same
// arithmetic) for the rest of the record, so no emission state
// stays live across the steps. The result is that of transport_photon bit
// for bit.
// ---------------------------------------------------------------------------
// Jobs per wave: a batch of up to 64 runs at the start of a step whenever 64
// or more are pending or the step's new jobs would not fit, so at most kQueue
// are pending after any step.
constexpr uint32_t kQueue = 96;
constexpr uint32_t kWaveQueueBytes = kQueue * 32 + 64 * 8;
constexpr int32_t kDeferredMaxSteps = (1 << 26) - 1;  // step index of a job: 26 bits

struct WaveQueue {
    float4* job;               // 2 per job: (position, step length), (direction, lane | step << 6)
    unsigned long long* key;   // per lane: earliest (step << 16 | target) hit, ~0 if none
};

__device__ __forceinline__ WaveQueue wave_queue(const KParams& p) {
    extern __shared__ float4 lds[];
    char* base = reinterpret_cast<char*>(lds) + p.queue_off + (threadIdx.x >> 6) * kWaveQueueBytes;
    WaveQueue q;
    q.job = reinterpret_cast<float4*>(base);
    q.key = reinterpret_cast<unsigned long long*>(base + kQueue * 32);
    return q;
}

// Orders the wave's LDS accesses: the LDS executes one wave's accesses in
// issue order, so only the compiler must not move them across this point.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One photon per lane (active lanes), the whole wave together.
template <bool COUNT, typename Search>
__device__ __forceinline__ bool transport_group(const KParams& p, const pp_source& src,
                                                uint32_t src_ix, uint64_t pid, bool active,
                                                const float* spec, const Search& search,
                                                const pp::Phase& phase, const WaveQueue& q,
                                                pp::Work& work, unsigned long long& n_steps_done,
                                                HitRec& h) {
    const uint32_t lane = __lane_id();
    pp::Stream s{p.k0, p.k1, (uint32_t)pid, (uint32_t)(pid >> 32), src_ix};
    const pp::StepStream ss(s);
    // the wave's photon ids are consecutive: pid = wave_first + lane
    const uint64_t pid0 = pid - lane;
    const uint64_t wave_first =
        ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(pid0 >> 32)) << 32) |
        __builtin_amdgcn_readfirstlane((uint32_t)pid0);
    f3 pos = pp::make3(0.0f, 0.0f, 0.0f), dir = pos;
    float ln2_sca = 0.0f;
    if (active) {
        // only what the steps need stays live; a hit emits the photon again
        const PhotonInit st = init_photon(p, src, spec, s);
        pos = st.pos;
        dir = st.dir;
        ln2_sca = st.mv.sca_len * pp::kLn2;
    }
    // no hit yet: all ones, materialised here (a 64-bit constant kept in a
    // register pair across the photon loop was spilled and reloaded per photon)
    uint32_t ones;
    __asm__ volatile("v_mov_b32 %0, -1" : "=v"(ones));
    *reinterpret_cast<uint2*>(q.key + lane) = make_uint2(ones, ones);
    // the stepping lanes as a wave mask (scalar). A photon whose hit is known
    // keeps stepping: its later steps cannot change the earliest (step,
    // target), and with ~1e-3 of the photons hit, stopping them would cost a
    // read of every lane's key after each batch (an LDS round trip) for
    // nothing
    const uint64_t going_m = __ballot(active);
    uint32_t head = 0, count = 0;  // wave-uniform ring of pending jobs
    // every lane walks one pending job (lanes >= n none)
    auto walk_jobs = [&](uint32_t n) {
        wave_sync();
        int found = -1;
        float dt = 0.0f;
        unsigned long long key = 0;
        uint32_t owner = 0;
        if (lane < n) {
            uint32_t slot = head + lane;
            if (slot >= kQueue) slot -= kQueue;
            const float4 a = q.job[2 * slot], b = q.job[2 * slot + 1];
            const uint32_t meta = __float_as_uint(b.w);
            owner = meta & 63u;
            if (COUNT) work.walks += 1;
            found = search.template find<COUNT>(pp::make3(a.x, a.y, a.z), pp::make3(b.x, b.y, b.z),
                                                a.w, dt, work);
            if (found >= 0) {
                key = ((unsigned long long)(meta >> 6) << 16) | (uint32_t)found;
                atomicMin(q.key + owner, key);
            }
        }
        head += n;
        if (head >= kQueue) head -= kQueue;
        count -= n;
    };
    // one loop for the steps and the final batches, so that the walk is
    // inlined once
    // The lane conditions below are these scalar masks, applied with
    // inverse ballots: no per-lane copy of a condition is kept or converted.
    int32_t step = 0;
    bool stepping = true;  // wave-uniform
    for (;;) {
        stepping = stepping && step < p.n_steps && going_m != 0;
        if (!stepping && count == 0) break;
        // r, step_size and next are set and read only while stepping (a
        // uniform branch); every lane computes them, the moving lanes
        // (move_m) use them (no register is set for the others)
        uint64_t move_m = 0, m = 0;
        uint4 r;
        float step_size;
        f3 next;
        if (stepping) {
            move_m = going_m;
            ++step;
            r = ss.block((uint32_t)step);
            step_size = pp::step_length(pp::u01_open0(r.x), ln2_sca);
            next = add_scaled(pos, step_size, dir);
            // a walk only for a segment near the grid; the walk clips it exactly
            m = search.touch_mask(pos, next) & move_m;
        }
        const uint32_t n_new = (uint32_t)__popcll(m);
        if (!stepping || count >= 64 || count + n_new > kQueue) walk_jobs(count < 64 ? count : 64);
        if (__builtin_amdgcn_inverse_ballot_w64(m)) {
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            uint32_t slot = head + count + rank;
            if (slot >= kQueue) slot -= kQueue;
            q.job[2 * slot] = make_float4(pos.x, pos.y, pos.z, step_size);
            q.job[2 * slot + 1] =
                make_float4(dir.x, dir.y, dir.z, __uint_as_float(lane | ((uint32_t)step << 6)));
        }
        count += n_new;
        if (__builtin_amdgcn_inverse_ballot_w64(move_m)) {
            pos = next;
            const float ct = pp::phase_cos_words(phase, r.y, r.z);
            dir = pp::update_direction(dir, ct, pp::u01(r.w));
        }
    }
    wave_sync();
    const unsigned long long key = q.key[lane];
    const bool hit = active && key != ~0ull;
    if (COUNT && active) n_steps_done += hit ? (unsigned long long)(key >> 16) : (unsigned long long)p.n_steps;
    if (hit) {
        // the record holds the photon id, module and step; replay_hits_kernel
        // completes it (emission again, steps 1..k-1, the sphere test of step
        // k for the distance within it), so that neither
        // the emission's state nor a second inlined emission stays in this
        // kernel's registers. The photon id is recomputed from the wave's first
        // id (scalar) and the lane.
        uint64_t first = wave_first;
        __asm__ volatile("" : "+s"(first));
        const uint64_t id = first + lane;
        const f3 zero = pp::make3(0.0f, 0.0f, 0.0f);
        h = HitRec{zero, zero, zero, 0.0f, 0.0f, 0.0f, (int32_t)(key & 0xFFFFu) + 1,
                   (int32_t)(key >> 16), id};
    }
    return hit;
}

// n_done: this lane's completed photons; wave_done: the wave's (the same
// value in every lane, added once)
template <bool COUNT>
__device__ __forceinline__ void flush_counts(const KParams& p, unsigned long long n_done,
                                             unsigned long long wave_done,
                                             unsigned long long n_steps_done, const pp::Work& work) {
    if (n_done) atomicAdd(&p.counters->n_ph_sim, n_done);
    if (wave_done && __lane_id() == 0) atomicAdd(&p.counters->n_ph_sim, wave_done);
    if (COUNT) {
        atomicAdd(&p.counters->steps, n_steps_done);
        atomicAdd(&p.counters->tests, (unsigned long long)work.tests);
        atomicAdd(&p.counters->cells, (unsigned long long)work.cells);
        atomicAdd(&p.counters->walks, (unsigned long long)work.walks);
    }
}

// Propagation of one source: cuda_propagate_photons! (src/photon_prop_cuda.jl:
// 429-549) with photons grid-strided over the launch.
// COUNT: instrumented variant that also counts photon steps, sphere tests
// and grid-cell visits (roofline accounting); the results are identical.
// Minimum waves per SIMD the register allocation must allow (the VGPR budget
// is 512 / waves per lane).
#ifndef PP_WAVES_PER_EU
#define PP_WAVES_PER_EU 6
#endif
#define PP_PROPAGATE_ATTRS \
    __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(PP_WAVES_PER_EU)))

// *p read through the constant address space: a scalar load into SGPRs for
// a wave-uniform address (the tables the kernels read are not written
// during a launch)
template <typename T>
__device__ __forceinline__ T load_constant(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(4))) T*)p;
#else
    return *p;
#endif
}

// The source record at sp, read through the constant address space (scalar
// loads into SGPRs), and the source index, both opaque to the optimiser: a
// caller that reads them in every iteration of a photon loop gets what the
// emission derives from the source computed in that iteration, not hoisted
// out of the loop into registers that spill to scratch.
__device__ __forceinline__ pp_source source_record(const pp_source* sp, uint32_t& si) {
#if defined(__HIP_DEVICE_COMPILE__)
    using ConstSource = const __attribute__((address_space(4))) pp_source;
    ConstSource* c = (ConstSource*)sp;
    __asm__ volatile("" : "+s"(c), "+s"(si));
    return *c;
#else
    return *sp;
#endif
}

template <int PHASE, int SEARCH, bool COUNT = false>
__global__ void PP_PROPAGATE_ATTRS propagate_kernel(KParams p) {
    const Lds l = stage_lds<SEARCH>(p);
    pp::Phase phase = p.med.phase;
    phase.kind = PHASE;
    const auto search = make_search<SEARCH>(p, l);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // photons completed: per lane in the one-photon loop, per wave (a scalar,
    // never spilled) in the wave loop
    unsigned long long n_done = 0, wave_done = 0, n_steps_done = 0;
    pp::Work work;
    if constexpr (kDeferredWalk<SEARCH>) {
        // whole waves of photons, the loop condition uniform across the wave
        const WaveQueue q = wave_queue(p);
        bool overflow = false;
        for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
             base < p.n_photons; base += stride) {
            uint32_t si = p.src_ix;
            const pp_source src = source_record(p.sources + p.src_ix, si);
            const uint64_t pid = base + __lane_id();
            const bool active = pid < p.n_photons && !overflow;
            HitRec h;
            const bool hit = transport_group<COUNT>(p, src, si, pid, active, l.spec, search,
                                                    phase, q, work, n_steps_done, h);
            const bool stored = store_hits(hit, h, si, p.hits, p.counters);
            if (!stored) {
                atomicOr(&p.counters->err, 1u);
                overflow = true;
            }
            wave_done += (unsigned long long)__popcll(__ballot(stored && active));
            if (__ballot(overflow)) break;
        }
    } else {
        for (uint64_t pid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; pid < p.n_photons;
             pid += stride) {
            uint32_t si = p.src_ix;
            const pp_source src = source_record(p.sources + p.src_ix, si);
            HitRec h;
            const bool hit = transport_photon<COUNT>(p, src, si, pid, l.spec, search,
                                                     phase, work, n_steps_done, h);
            if (!store_hits(hit, h, si, p.hits, p.counters)) {
                atomicOr(&p.counters->err, 1u);
                break;
            }
            ++n_done;
        }
    }
    flush_counts<COUNT>(p, n_done, wave_done, n_steps_done, work);
}

// Batch of sources in one persistent launch. Work items are chunks of
// `chunk` consecutive photons of one source (chunk_prefix[s] = first chunk
// of source s). Each wave takes the next work item from a global counter
// (work stealing: waves that drew cheap photons simply take more chunks)
// and exits when the counter passes the last item, so every wave ends.
// Chunks from split_from on are handed out as `parts` pieces of `piece`
// photons each (the last piece of a chunk may be shorter), so that the waves
// end closer together. Photon ids and random streams are those of the
// one-launch-per-source kernel: the hits are identical.
struct BatchParams {
    const pp_source* sources;
    const uint64_t* n_photons;     // per source
    const uint64_t* chunk_prefix;  // n_sources + 1
    int32_t n_sources;
    uint32_t chunk;                // multiple of 64
    uint32_t parts, piece;         // pieces per split chunk; photons per piece, a multiple of 64
    uint64_t split_from;           // first split chunk (chunk_prefix[n_sources]: none)
};

template <int PHASE, int SEARCH, bool COUNT = false>
__global__ void PP_PROPAGATE_ATTRS propagate_batch_kernel(KParams p, BatchParams b) {
    const Lds l = stage_lds<SEARCH>(p);
    pp::Phase phase = p.med.phase;
    phase.kind = PHASE;
    const auto search = make_search<SEARCH>(p, l);
    const uint64_t n_chunks = load_constant(b.chunk_prefix + b.n_sources);
    const uint64_t total = b.split_from + (n_chunks - b.split_from) * b.parts;
    const uint32_t lane = __lane_id();
    WaveQueue q{};
    if constexpr (kDeferredWalk<SEARCH>) q = wave_queue(p);
    unsigned long long n_done = 0, wave_done = 0, n_steps_done = 0;
    pp::Work work;
    bool overflow = false;
    for (;;) {
        unsigned long long got = 0;
        if (lane == 0) got = atomicAdd(&p.counters->next_chunk, 1ull);
        const uint64_t ck = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(got >> 32)) << 32) |
                            __builtin_amdgcn_readfirstlane((uint32_t)got);
        if (ck >= total) break;
        // chunk c, photons [off, off + len) of it
        uint64_t c = ck;
        uint32_t off = 0, len = b.chunk;
        if (ck >= b.split_from) {
            const uint32_t j = (uint32_t)(ck - b.split_from);
            c = b.split_from + j / b.parts;
            off = (j % b.parts) * b.piece;
            len = b.chunk - off < b.piece ? b.chunk - off : b.piece;
        }
        // source of this chunk: last s with chunk_prefix[s] <= c
        int lo = 0, hi = b.n_sources - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (load_constant(b.chunk_prefix + mid) <= c) lo = mid;
            else hi = mid - 1;
        }
        const int si0 = lo;
        const uint64_t n_src = load_constant(b.n_photons + si0);
        const uint64_t first = (c - load_constant(b.chunk_prefix + si0)) * b.chunk + off;
        const uint64_t end = first + len < n_src ? first + len : n_src;
        for (uint64_t base = first; base < end; base += 64) {
            // the source index and record, opaque to the optimiser in every
            // iteration: what the emission derives from them is computed here
            // (one scalar reload per 64 photons) instead of being hoisted out
            // of the loop into registers that then spill to scratch
            uint32_t si = (uint32_t)si0;
            const pp_source src = source_record(b.sources + si0, si);
            const uint64_t pid = base + lane;
            if constexpr (kDeferredWalk<SEARCH>) {
                const bool active = pid < end && !overflow;
                HitRec h;
                const bool hit = transport_group<COUNT>(p, src, (uint32_t)si, pid, active, l.spec,
                                                        search, phase, q, work, n_steps_done, h);
                const bool stored = store_hits(hit, h, (uint32_t)si, p.hits, p.counters);
                if (!stored) {
                    atomicOr(&p.counters->err, 1u);
                    overflow = true;
                }
                wave_done += (unsigned long long)__popcll(__ballot(stored && active));
            } else {
                if (pid >= end || overflow) continue;
                HitRec h;
                const bool hit = transport_photon<COUNT>(p, src, (uint32_t)si, pid, l.spec, search,
                                                         phase, work, n_steps_done, h);
                if (!store_hits(hit, h, (uint32_t)si, p.hits, p.counters)) {
                    atomicOr(&p.counters->err, 1u);
                    overflow = true;
                    continue;
                }
                ++n_done;
            }
        }
        if (__ballot(overflow)) break;  // the call fails; stop taking chunks
    }
    flush_counts<COUNT>(p, n_done, wave_done, n_steps_done, work);
}

// Completion of the hit records of the deferred-walk kernels (transport_group):
// a record holds the photon id, module (target + 1), step k and source. One
// lane per record emits the photon again, replays steps 1..k-1 with
// transport_photon's arithmetic on the same random stream (contraction off:
// the same bits as the one-photon schedule), takes step k's distance to the
// target from the walk's own sphere test on the same operands, and writes
// position, directions, wavelength, time and distance.
// sources: the call's source list on the device.
__global__ void __launch_bounds__(256) replay_hits_kernel(KParams p, const pp_source* sources) {
    const pp::Phase phase = p.med.phase;
    const unsigned long long n_hits = p.counters->n_hits;
    const uint64_t n = n_hits < (unsigned long long)p.hits.capacity ? n_hits : p.hits.capacity;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float4* r = reinterpret_cast<float4*>(p.hits.rec + i);
        const float4 r3 = r[3];
        const uint64_t id = (uint64_t)__float_as_uint(r3.x) | ((uint64_t)__float_as_uint(r3.y) << 32);
        const uint64_t pid = id & (kMaxPhotonsPerSource - 1);
        const int32_t target = (int32_t)((id >> kPidBits) & 0xFFFFu) - 1;
        const int32_t k = __float_as_int(r3.z);
        const uint32_t si = __float_as_uint(r3.w);
        const pp_source src = sources[si];
        const pp::Stream s{p.k0, p.k1, (uint32_t)pid, (uint32_t)(pid >> 32), si};
        const pp::StepStream ss(s);
        const PhotonInit st = init_photon(p, src, p.spec, s);
        const float ln2_sca = st.mv.sca_len * pp::kLn2, inv_c = st.mv.inv_c_grp;
        f3 rp = st.pos, rd = st.dir;
        float rt = st.time, rdist = 0.0f;
        for (int32_t j = 1; j < k; ++j) {
            const uint4 w = ss.block((uint32_t)j);
            const float step_size = pp::step_length(pp::u01_open0(w.x), ln2_sca);
            move_by(rp, rdist, rt, rd, step_size, inv_c);
            const float ct = pp::phase_cos_words(phase, w.y, w.z);
            rd = pp::update_direction(rd, ct, pp::u01(w.w));
        }
        // step k, cut at the target: the distance as GridSearch's walk finds it
        const uint4 w = ss.block((uint32_t)k);
        const float step_size = pp::step_length(pp::u01_open0(w.x), ln2_sca);
        float d = 0.0f;
        pp::sphere_hit(p.targets[target], rp, rd, pp::dir_excess(rd), step_size, d);
        move_by(rp, rdist, rt, rd, d, inv_c);
        r[0] = make_float4(rp.x, rp.y, rp.z, rd.x);
        r[1] = make_float4(rd.y, rd.z, st.dir.x, st.dir.y);
        r[2] = make_float4(st.dir.z, st.wl, rt, rdist);
    }
}

// Zeroes the counters in a captured graph (a kernel node, not a memset node).
__global__ void clear_counters_kernel(Counters* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = Counters{};
}

// ---------------------------------------------------------------------------
// step-function evaluation kernel (parity tests)
// ---------------------------------------------------------------------------
struct EvalParams {
    int fn;
    int in_stride, out_stride;
    int64_t n;
    const float* in;
    float* out;
    pp::Medium med;
    const float* spec;
    int spec_n, spec_map;
    float spec_mono;
    float spec_nf, spec_lastf;  // float(spec_n), float(spec_n - 1)
};

__device__ __forceinline__ uint32_t bits(float x) { return __float_as_uint(x); }
__device__ __forceinline__ float as_f(uint32_t x) { return __uint_as_float(x); }

__global__ void __launch_bounds__(256) eval_fn_kernel(EvalParams e) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e.n) return;
    const float* x = e.in + i * e.in_stride;
    float* y = e.out + i * e.out_stride;
    switch (e.fn) {
    case PP_FN_ISOTROPIC_DIR: {
        f3 d = pp::isotropic_dir(x[0], x[1]);
        y[0] = d.x; y[1] = d.y; y[2] = d.z;
        break;
    }
    case PP_FN_ROT_FROM_EZ: {
        f3 d = pp::rot_from_ez(pp::make3(x[0], x[1], x[2]), pp::make3(x[3], x[4], x[5]));
        y[0] = d.x; y[1] = d.y; y[2] = d.z;
        break;
    }
    case PP_FN_HG_COS: {
        pp::Phase ph = pp::make_phase(PP_PHASE_HG, x[0], 1.0f, 1.0f);
        y[0] = pp::hg_cos(ph, x[1]);
        break;
    }
    case PP_FN_ES_COS: {
        pp::Phase ph = pp::make_phase(PP_PHASE_MIXED_HG_ES, 0.5f, x[0], 0.0f);
        y[0] = pp::es_cos(ph, x[1]);
        break;
    }
    case PP_FN_UPDATE_DIR: {
        f3 d = pp::update_direction(pp::make3(x[0], x[1], x[2]), x[3], x[4]);
        y[0] = d.x; y[1] = d.y; y[2] = d.z;
        break;
    }
    case PP_FN_SPHERE_ISEC: {
        float d;
        const pp::f3 dir = pp::make3(x[7], x[8], x[9]);
        const bool h = pp::sphere_hit(make_float4(x[0], x[1], x[2], x[3] * x[3]),
                                      pp::make3(x[4], x[5], x[6]), dir, pp::dir_excess(dir),
                                      x[10], d);
        y[0] = h ? 1.0f : 0.0f;
        y[1] = h ? d : 0.0f;
        break;
    }
    case PP_FN_STEP_LEN:
        y[0] = pp::step_length(x[0], x[1] * pp::kLn2);
        break;
    case PP_FN_CASCADIA: {
        pp::MediumValues mv = pp::medium_values(e.med, x[0]);
        y[0] = mv.sca_len;
        y[1] = 1.0f / mv.inv_c_grp;
        y[2] = 1.0f / mv.cos_c;
        break;
    }
    case PP_FN_TRACK_DIR: {
        f3 d = pp::cherenkov_track_dir(x[0], x[1]);
        y[0] = d.x; y[1] = d.y; y[2] = d.z;
        break;
    }
    case PP_FN_PHILOX: {
        // the step path's StepStream, Stream{k0, k1, c1, c2, c3}.block(c0), and
        // the plain ten rounds of the emission draws: all ones where they differ
        const pp::Stream s{bits(x[4]), bits(x[5]), bits(x[1]), bits(x[2]), bits(x[3])};
        uint4 r = pp::StepStream(s).block(bits(x[0]));
        const uint4 g = pp::philox(make_uint4(bits(x[0]), bits(x[1]), bits(x[2]), bits(x[3])),
                                   bits(x[4]), bits(x[5]));
        if (r.x != g.x || r.y != g.y || r.z != g.z || r.w != g.w) r = make_uint4(~0u, ~0u, ~0u, ~0u);
        y[0] = as_f(r.x); y[1] = as_f(r.y); y[2] = as_f(r.z); y[3] = as_f(r.w);
        break;
    }
    case PP_FN_GAMMA: {
        pp::Stream s{bits(x[2]), bits(x[3]), bits(x[4]), bits(x[5]), bits(x[6])};
        float xa = 0.0f;
        y[0] = pp::gamma_sample(x[0], x[1], s, PP_SUB_LONG_GAMMA, &xa);
        y[1] = xa;
        break;
    }
    case PP_FN_SPECTRUM:
        y[0] = pp::spectrum_sample(e.spec, e.spec_n, e.spec_nf, e.spec_lastf, e.spec_map,
                                   e.spec_mono, x[0]);
        break;
    case PP_FN_SL_COS: {
        pp::Phase ph = pp::make_phase(PP_PHASE_MIXED_HG_SL, x[0], 1.0f, 0.0f);
        y[0] = pp::sl_cos(ph, x[1]);
        break;
    }
    case PP_FN_MIXED_HGES: {
        // the sampler the kernels run for a MixedHGES medium, on the random
        // words whose u01 are x[3], x[4] (multiples of 2^-24 in [0, 1))
        pp::Phase ph = pp::make_phase(PP_PHASE_MIXED_HG_ES, x[0], x[1], x[2]);
        const uint32_t wc = (uint32_t)(x[3] * 0x1p24f) << 8, wu = (uint32_t)(x[4] * 0x1p24f) << 8;
        y[0] = pp::phase_cos_words(ph, wc, wu);
        // the words form must give the float form's bits: NaN where it does not
        if (__float_as_uint(y[0]) != __float_as_uint(pp::hg_es_cos(ph, pp::u01(wc), pp::u01(wu))))
            y[0] = __uint_as_float(0x7fc00000u);
        break;
    }
    case PP_FN_PLANE_ISEC: {
        float d;
        const uint32_t kind = bits(x[0]);
        const float4 t = make_float4(x[1], x[2], x[3], x[4] * x[4]);
        const float4 ext = make_float4(x[0], x[4] / 2.0f, x[5] / 2.0f, 0.0f);
        const bool h = pp::plane_hit(kind, t, ext, pp::make3(x[6], x[7], x[8]),
                                     pp::make3(x[9], x[10], x[11]), x[12], d);
        y[0] = h ? 1.0f : 0.0f;
        y[1] = h ? d : 0.0f;
        break;
    }
    default:
        break;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define PP_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(PP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

pp::Medium lower_medium(const pp_medium& m) {
    pp::Medium d{};
    d.kind = m.kind;
    d.phase = pp::make_phase(m.phase_kind, m.hg_g, m.es_b, m.hg_fraction);
    d.sca_len = m.sca_len;
    d.c_grp = m.group_velocity;
    d.inv_c_grp = 1.0f / m.group_velocity;
    d.n_phase = m.phase_refractive_index;
    d.sca_scale = (double)m.sca_scale;
    d.kop1 = (double)0.0017f;
    d.kop2 = (double)(float)(1.34 * (double)m.vol_conc_small);
    d.kop3 = (double)(float)(0.312 * (double)m.vol_conc_large);
    // Quan & Fry (1995) coefficients; linear pressure term 1.45e-5 / atm
    const double S = m.salinity, T = m.temperature, P = m.pressure;
    d.qf0 = (float)(1.31405 + (1.779e-4 - 1.05e-6 * T + 1.6e-8 * T * T) * S - 2.02e-6 * T * T +
                    1.45e-5 * P);
    d.qf1 = (float)(15.868 + 0.01155 * S - 0.00423 * T);
    d.qf2 = (float)(-4382.0);
    d.qf3 = (float)(1.1455e6);
    d.qg1 = 2.0f * d.qf1;
    d.qg2 = 3.0f * d.qf2;
    d.qg3 = 4.0f * d.qf3;
    if (m.kind == PP_MEDIUM_ICE) {
        d.ice_pref = (float)((double)m.sca_scale * (1.0 - (double)m.hg_g) / (double)m.ice_b_dust_400);
        d.ice_alpha = m.ice_alpha_sca_dust;
    }
    return d;
}

bool finite3(const float* v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

}  // namespace

struct pp_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    pp_config cfg{};
    void* hit_mem = nullptr;
    HitCols hits{};
    Counters* d_counters = nullptr;
    bool has_medium = false;
    pp_medium medium_in{};
    pp::Medium medium{};
    bool has_spectrum = false;
    float spec_mono = 0.0f;
    int spec_n = 0, spec_map = 0;
    float* d_spec = nullptr;
    int spec_cap = 0;
    int n_targets = -1;
    float4* d_targets = nullptr;
    int tgt_cap = 0;
    std::vector<float4> h_targets;  // x, y, z, r^2
    std::vector<float> h_radius;
    bool planar = false;  // the list holds rectangles or discs: two records per target
    // uniform grid over the targets, rebuilt when targets or cell size change
    pp::GridDesc grid{};
    float grid_cell_built = -1.0f;  // requested cell size of the built grid
    bool grid_global = false;        // the built grid's tables exceed the LDS: read from memory
    float grid_cell_z_built = 0.0f;
    float4* d_cells = nullptr;
    uint16_t* d_cell_items = nullptr;
    size_t cell_cap = 0, item_cap = 0;
    int last_search = 0;
    int64_t n_hits = 0, n_ph_sim = 0;
    int64_t work_steps = 0, work_tests = 0, work_cells = 0, work_walks = 0;
    float last_ms = 0.0f;
    int32_t n_launch = 0;
    int block = 256;  // threads per block of the last launch (cfg.block_size, or chosen for 0)
    bool perm_valid = false;
    std::vector<int64_t> perm;
    // pinned host staging of hit records for pp_copy_hits
    HitRecord* h_rec = nullptr;
    int64_t h_rec_cap = 0;
    // device sort of the hits by (source, photon id): the records in that
    // order, keys, indices and the radix sort's scratch, in one allocation
    char* d_sort = nullptr;
    size_t d_sort_cap = 0;
    bool sorted_valid = false;  // d_sort holds the sorted records of the last propagation
    int32_t last_n_sources = 0;
    // source table of the persistent batch kernel: device copy and pinned
    // host staging, laid out sources | photon counts | chunk prefix
    void* d_batch = nullptr;
    void* h_batch = nullptr;
    size_t batch_cap = 0;  // bytes of each
    // PMT detection of the context's hits: targets as (x, y, z, radius), and
    // one device block holding the acceptance table | detected count | PMT ids
    float4* d_tgt_r = nullptr;
    char* d_pmt = nullptr;
    size_t pmt_cap = 0;
    // hipGraph of the batch mode PP_BATCH_GRAPH: counter clear -> propagation
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraphNode_t graph_prop = nullptr;
    const void* graph_fn = nullptr;
    unsigned graph_blocks = 0, graph_block = 0;
    size_t graph_lds = 0;
};

namespace {

// Make the context's device current and drop any error state left by an
// earlier, already reported call, so that launch checks see only their own.
int set_dev(pp_ctx* c) {
    PP_HIP(hipSetDevice(c->device));
    (void)hipGetLastError();
    return PP_OK;
}

constexpr int kMaxLdsBytes = 160 * 1024;
// records per pinned staging chunk of pp_copy_hits (64 MB)
constexpr int64_t kStageRecords = (int64_t)1 << 20;
constexpr size_t kGridLdsBudget = 64 * 1024;  // grid tables + targets + spectrum per block
// the largest table set that leaves room for the walk queues of 16 waves
constexpr size_t kGridLdsMax = kMaxLdsBytes - 16 * (size_t)kWaveQueueBytes;
constexpr size_t kGridGlobalBudget = (size_t)64 << 20;  // tables read from device memory

// (Re)build the target grid for the requested cell edge and upload it.
int build_grid(pp_ctx* c, float cell_req, float cell_z_req) {
    std::vector<float> xyz(3 * (size_t)c->n_targets);
    for (int i = 0; i < c->n_targets; ++i) {
        xyz[3 * i] = c->h_targets[i].x;
        xyz[3 * i + 1] = c->h_targets[i].y;
        xyz[3 * i + 2] = c->h_targets[i].z;
    }
    pp::GridTables t;
    // the occupancy-friendly budget first; for many targets (their records
    // alone may fill it) the whole LDS, at fewer workgroups per CU; beyond
    // that, tables in device memory that the search reads through the caches
    bool global = false;
    if (!pp::build_grid_tables(xyz.data(), c->h_radius.data(), c->n_targets, cell_req, cell_z_req,
                               c->spec_n, kGridLdsBudget, t) &&
        !pp::build_grid_tables(xyz.data(), c->h_radius.data(), c->n_targets, cell_req, cell_z_req,
                               c->spec_n, kGridLdsMax, t)) {
        global = true;
        if (!pp::build_grid_tables(xyz.data(), c->h_radius.data(), c->n_targets, cell_req,
                                   cell_z_req, c->spec_n, kGridGlobalBudget, t))
            return fail(PP_ERR_INVALID, "grid: no cell size fits");
    }
    if (t.cells.size() > c->cell_cap) {
        if (c->d_cells) PP_HIP(hipFree(c->d_cells));
        c->d_cells = nullptr;
        PP_HIP(hipMalloc(&c->d_cells, t.cells.size() * sizeof(float4)));
        c->cell_cap = t.cells.size();
    }
    if (t.items.size() > c->item_cap) {
        if (c->d_cell_items) PP_HIP(hipFree(c->d_cell_items));
        c->d_cell_items = nullptr;
        PP_HIP(hipMalloc(&c->d_cell_items, t.items.size() * 2));
        c->item_cap = t.items.size();
    }
    // uploads are ordered on the context stream, where the kernels that read
    // them run (a blocking-stream copy from pageable memory may still be in
    // flight when it returns); the sync keeps the host tables alive until done
    PP_HIP(hipMemcpyAsync(c->d_cells, t.cells.data(), t.cells.size() * sizeof(float4),
                          hipMemcpyHostToDevice, c->stream));
    PP_HIP(hipMemcpyAsync(c->d_cell_items, t.items.data(), t.items.size() * 2,
                          hipMemcpyHostToDevice, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    c->grid = t.g;
    c->grid_global = global;
    c->grid_cell_built = cell_req;
    c->grid_cell_z_built = cell_z_req;
    return PP_OK;
}

// Defaults and range checks of a pp_config.
int normalise_config(pp_config& c, const char* who) {
    const std::string w(who);
    // block_size 0: chosen per call from the LDS the launch needs (pp_propagate)
    if (c.block_size != 0 && (c.block_size < 64 || c.block_size > 1024 || c.block_size % 64))
        return fail(PP_ERR_INVALID, w + ": block_size must be 0 or a multiple of 64 in [64,1024]");
    if (c.waves_per_simd < 0 || c.waves_per_simd > 32)
        return fail(PP_ERR_INVALID, w + ": waves_per_simd must be in [0,32]");
    if (c.search < PP_SEARCH_AUTO || c.search > PP_SEARCH_GRID)
        return fail(PP_ERR_INVALID, w + ": unknown search");
    if (!(c.grid_cell >= 0.0f) || !std::isfinite(c.grid_cell))
        return fail(PP_ERR_INVALID, w + ": grid_cell must be finite and >= 0");
    if (!std::isfinite(c.grid_cell_z))
        return fail(PP_ERR_INVALID, w + ": grid_cell_z must be finite");
    if (c.batch_mode < PP_BATCH_AUTO || c.batch_mode > PP_BATCH_GRAPH)
        return fail(PP_ERR_INVALID, w + ": unknown batch_mode");
    if (c.chunk_photons < 0 || c.chunk_photons % 64 || c.chunk_photons > (1 << 24))
        return fail(PP_ERR_INVALID, w + ": chunk_photons must be a multiple of 64 in [0, 2^24]");
    if (c.deterministic_order < 0 || c.deterministic_order > 2)
        return fail(PP_ERR_INVALID, w + ": deterministic_order must be 0, 1 or 2");
    return PP_OK;
}

// Blocks that fill the device: occupancy limit or cfg.waves_per_simd.
int max_resident_blocks(pp_ctx* c, const void* fn, int block, size_t lds, int64_t& out) {
    int occ = 0;
    PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, lds));
    // LDS is allocated in 512-byte units per workgroup, which the query leaves out
    if (lds > 0) occ = std::min<int>(occ, (int)((size_t)kMaxLdsBytes / ((lds + 511) & ~(size_t)511)));
    if (occ < 1) return fail(PP_ERR_INVALID, "pp_propagate: kernel does not fit one block per CU");
    out = (int64_t)occ * c->n_cu;
    if (c->cfg.waves_per_simd > 0)
        out = std::max<int64_t>(1, (int64_t)c->n_cu * 4 * c->cfg.waves_per_simd * 64 / block);
    return PP_OK;
}

// The context's source table (device) and its pinned staging, grown to at
// least `bytes`; waits for the stream first, since earlier work may read them.
int ensure_batch_table(pp_ctx* c, size_t bytes) {
    PP_HIP(hipStreamSynchronize(c->stream));
    if (bytes <= c->batch_cap) return PP_OK;
    if (c->d_batch) PP_HIP(hipFree(c->d_batch));
    if (c->h_batch) PP_HIP(hipHostFree(c->h_batch));
    c->d_batch = c->h_batch = nullptr;
    c->batch_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 64 * 1024);
    PP_HIP(hipMalloc(&c->d_batch, cap));
    PP_HIP(hipHostMalloc(&c->h_batch, cap, hipHostMallocDefault));
    c->batch_cap = cap;
    return PP_OK;
}

// The persistent kernel of a search kind and phase: the one whose occupancy
// chooses the automatic block size.
const void* batch_kernel_fn(int skind, int phase) {
    using BatchFn = void (*)(KParams, BatchParams);
#define PP_K(PH, S) propagate_batch_kernel<PH, S, false>
#define PP_K3(S) {PP_K(PP_PHASE_HG, S), PP_K(PP_PHASE_MIXED_HG_ES, S), PP_K(PP_PHASE_MIXED_HG_SL, S)}
    static const BatchFn kKernels[4][3] = {PP_K3(kSearchBrute), PP_K3(kSearchGrid),
                                           PP_K3(kSearchMixed), PP_K3(kSearchGridGlobal)};
#undef PP_K3
#undef PP_K
    return (const void*)kKernels[skind][phase];
}

// One launch per source, as src/photon_prop_cuda.jl:925-944.
int launch_per_source(pp_ctx* c, KParams kp, const pp_source* sources,
                      const std::vector<uint64_t>& nph, int skind, int phase, bool count, size_t lds) {
    using KernelFn = void (*)(KParams);
#define PP_K(PH, S, C) propagate_kernel<PH, S, C>
#define PP_K3(S, C) {PP_K(PP_PHASE_HG, S, C), PP_K(PP_PHASE_MIXED_HG_ES, S, C), PP_K(PP_PHASE_MIXED_HG_SL, S, C)}
    // [count][search kind][phase kind]
    static const KernelFn kKernels[2][4][3] = {
        {PP_K3(kSearchBrute, false), PP_K3(kSearchGrid, false), PP_K3(kSearchMixed, false),
         PP_K3(kSearchGridGlobal, false)},
        {PP_K3(kSearchBrute, true), PP_K3(kSearchGrid, true), PP_K3(kSearchMixed, true),
         PP_K3(kSearchGridGlobal, true)}};
#undef PP_K3
#undef PP_K
    const KernelFn kern = kKernels[count][skind][phase];
    const int block = c->block;
    int64_t max_blocks = 0;
    if (int rc = max_resident_blocks(c, (const void*)kern, block, lds, max_blocks)) return rc;
    // the source list on the device (the kernels read their source from it)
    const size_t bytes = nph.size() * sizeof(pp_source);
    if (int rc = ensure_batch_table(c, bytes)) return rc;
    std::memcpy(c->h_batch, sources, bytes);
    PP_HIP(hipMemcpyAsync(c->d_batch, c->h_batch, bytes, hipMemcpyHostToDevice, c->stream));
    kp.sources = (const pp_source*)c->d_batch;
    PP_HIP(hipMemsetAsync(c->d_counters, 0, sizeof(Counters), c->stream));
    PP_HIP(hipEventRecord(c->ev0, c->stream));
    for (size_t i = 0; i < nph.size(); ++i) {
        if (nph[i] == 0) continue;
        kp.src_ix = (uint32_t)i;
        kp.n_photons = nph[i];
        const int64_t need = (int64_t)((nph[i] + block - 1) / block);
        const int blocks = (int)std::min<int64_t>(need, max_blocks);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(block), lds, c->stream, kp);
        PP_HIP(hipGetLastError());
        ++c->n_launch;
    }
    return PP_OK;
}

// The whole source list in one persistent launch (propagate_batch_kernel);
// with use_graph, as a hipGraph of counter clear + propagation.
int launch_batch(pp_ctx* c, const KParams& kp, const pp_source* sources,
                 const std::vector<uint64_t>& nph, int skind, int phase, bool count, size_t lds,
                 bool use_graph) {
    using BatchFn = void (*)(KParams, BatchParams);
#define PP_K(PH, S, C) propagate_batch_kernel<PH, S, C>
#define PP_K3(S, C) {PP_K(PP_PHASE_HG, S, C), PP_K(PP_PHASE_MIXED_HG_ES, S, C), PP_K(PP_PHASE_MIXED_HG_SL, S, C)}
    static const BatchFn kKernels[2][4][3] = {
        {PP_K3(kSearchBrute, false), PP_K3(kSearchGrid, false), PP_K3(kSearchMixed, false),
         PP_K3(kSearchGridGlobal, false)},
        {PP_K3(kSearchBrute, true), PP_K3(kSearchGrid, true), PP_K3(kSearchMixed, true),
         PP_K3(kSearchGridGlobal, true)}};
#undef PP_K3
#undef PP_K
    const BatchFn kern = kKernels[count][skind][phase];
    const int block = c->block;
    const size_t n = nph.size();
    uint64_t total_ph = 0;
    for (uint64_t v : nph) total_ph += v;
    if (total_ph == 0) {
        // nothing to run; the counters still have to read zero
        PP_HIP(hipMemsetAsync(c->d_counters, 0, sizeof(Counters), c->stream));
        PP_HIP(hipEventRecord(c->ev0, c->stream));
        return PP_OK;
    }
    int64_t max_blocks = 0;
    if (int rc = max_resident_blocks(c, (const void*)kern, block, lds, max_blocks)) return rc;
    const int64_t waves = max_blocks * (block / 64);
    // chunk: about 32 chunks per resident wave, so that the last chunks
    // taken leave little tail, and at least 256 photons: every grab is an
    // atomic on one counter, and the device serves those at about one per
    // 12 ns (a 1e8-photon launch of 192-photon chunks took 6.7 ms on them
    // alone; 5.2 ms with 512)
    uint64_t chunk = (uint64_t)c->cfg.chunk_photons;
    if (chunk == 0) {
        chunk = total_ph / ((uint64_t)waves * 32) / 64 * 64;
        chunk = std::min<uint64_t>(std::max<uint64_t>(chunk, 256), 8192);
    }
    // table: sources | photon counts | chunk prefix, each 256-byte aligned
    const size_t off_n = ((n * sizeof(pp_source)) + 255) & ~(size_t)255;
    const size_t off_p = off_n + ((n * 8 + 255) & ~(size_t)255);
    const size_t bytes = off_p + (n + 1) * 8;
    if (int rc = ensure_batch_table(c, bytes)) return rc;
    char* h = (char*)c->h_batch;
    std::memcpy(h, sources, n * sizeof(pp_source));
    uint64_t* hn = (uint64_t*)(h + off_n);
    uint64_t* hp = (uint64_t*)(h + off_p);
    hp[0] = 0;
    for (size_t i = 0; i < n; ++i) {
        hn[i] = nph[i];
        hp[i + 1] = hp[i] + (nph[i] + chunk - 1) / chunk;
    }
    const uint64_t n_chunks = hp[n];
    char* d = (char*)c->d_batch;
    BatchParams bp{};
    bp.sources = (const pp_source*)d;
    bp.n_photons = (const uint64_t*)(d + off_n);
    bp.chunk_prefix = (const uint64_t*)(d + off_p);
    bp.n_sources = (int32_t)n;
    bp.chunk = (uint32_t)chunk;
    // With several sources the chunks' costs differ (a track's photons are
    // not a cascade's), and the last two chunks per resident wave go out in
    // pieces of 256 photons (the 256-source batch: 27.45 -> 27.05 ms). One
    // source keeps whole chunks: its tail is even, and the pieces' extra
    // grabs on the one counter cost more than they save.
    bp.piece = (uint32_t)chunk;
    bp.parts = 1;
    bp.split_from = n_chunks;
    if (n > 1) {
        bp.piece = (uint32_t)std::min<uint64_t>(chunk, 256);
        bp.parts = (uint32_t)((chunk + bp.piece - 1) / bp.piece);
        const uint64_t tail = (uint64_t)waves * 2;
        bp.split_from = n_chunks > tail ? n_chunks - tail : 0;
    }
    const uint64_t n_items = bp.split_from + (n_chunks - bp.split_from) * bp.parts;
    const int64_t want = (int64_t)((n_items + block / 64 - 1) / (block / 64));
    const unsigned blocks = (unsigned)std::min<int64_t>(want, max_blocks);
    PP_HIP(hipMemcpyAsync(c->d_batch, c->h_batch, bytes, hipMemcpyHostToDevice, c->stream));
    if (!use_graph) {
        PP_HIP(hipMemsetAsync(c->d_counters, 0, sizeof(Counters), c->stream));
        PP_HIP(hipEventRecord(c->ev0, c->stream));
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(block), lds, c->stream, kp, bp);
        PP_HIP(hipGetLastError());
        c->n_launch = 1;
        return PP_OK;
    }
    KParams kp_arg = kp;
    void* args[] = {&kp_arg, &bp};
    hipKernelNodeParams np{};
    np.func = (void*)kern;
    np.gridDim = dim3(blocks);
    np.blockDim = dim3(block);
    np.sharedMemBytes = (unsigned)lds;
    np.kernelParams = args;
    np.extra = nullptr;
    const bool same_shape = c->graph_exec && c->graph_fn == (const void*)kern &&
                            c->graph_blocks == blocks && c->graph_block == (unsigned)block &&
                            c->graph_lds == lds;
    bool updated = false;
    if (same_shape)
        updated = hipGraphExecKernelNodeSetParams(c->graph_exec, c->graph_prop, &np) == hipSuccess;
    if (!updated) {
        (void)hipGetLastError();
        if (c->graph_exec) PP_HIP(hipGraphExecDestroy(c->graph_exec));
        if (c->graph) PP_HIP(hipGraphDestroy(c->graph));
        c->graph_exec = nullptr;
        c->graph = nullptr;
        PP_HIP(hipGraphCreate(&c->graph, 0));
        // the counters are cleared by a kernel node, not a memset node
        Counters* ctr = c->d_counters;
        void* cargs[] = {&ctr};
        hipKernelNodeParams cp{};
        cp.func = (void*)clear_counters_kernel;
        cp.gridDim = dim3(1);
        cp.blockDim = dim3(64);
        cp.sharedMemBytes = 0;
        cp.kernelParams = cargs;
        cp.extra = nullptr;
        hipGraphNode_t clear = nullptr;
        PP_HIP(hipGraphAddKernelNode(&clear, c->graph, nullptr, 0, &cp));
        PP_HIP(hipGraphAddKernelNode(&c->graph_prop, c->graph, &clear, 1, &np));
        PP_HIP(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        c->graph_fn = (const void*)kern;
        c->graph_blocks = blocks;
        c->graph_block = (unsigned)block;
        c->graph_lds = lds;
    }
    PP_HIP(hipEventRecord(c->ev0, c->stream));
    PP_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    c->n_launch = 1;
    return PP_OK;
}

}  // namespace

extern "C" {

int pp_abi_version(void) { return PP_ABI_VERSION; }

const char* pp_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// used by pp_mlp.hip, the surrogate half of the library
void pp_set_error_message(const char* msg) { g_err = msg; }

extern "C" {

int pp_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pp_device_memory(int device, int64_t* total_bytes, int64_t* free_bytes) {
    int ndev = pp_device_count();
    if (device < 0 || device >= ndev) return fail(PP_ERR_INVALID, "pp_device_memory: no such device");
    PP_HIP(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    PP_HIP(hipMemGetInfo(&fr, &tot));
    if (total_bytes) *total_bytes = (int64_t)tot;
    if (free_bytes) *free_bytes = (int64_t)fr;
    return PP_OK;
}

int pp_device_info(int device, char* arch, int arch_len, int32_t* n_cu) {
    int ndev = pp_device_count();
    if (device < 0 || device >= ndev) return fail(PP_ERR_INVALID, "pp_device_info: no such device");
    hipDeviceProp_t prop{};
    PP_HIP(hipGetDeviceProperties(&prop, device));
    if (arch && arch_len > 0) {
        std::string a(prop.gcnArchName);
        a = a.substr(0, a.find(':'));
        std::snprintf(arch, (size_t)arch_len, "%s", a.c_str());
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    return PP_OK;
}

int64_t pp_hit_record_bytes(void) {
    return (int64_t)sizeof(HitRecord);
}

int pp_create(pp_ctx** out, int device, int64_t hit_capacity, const pp_config* cfg) {
    if (!out) return fail(PP_ERR_INVALID, "pp_create: out is NULL");
    *out = nullptr;
    if (hit_capacity <= 0) return fail(PP_ERR_INVALID, "pp_create: hit_capacity must be > 0");
    int ndev = pp_device_count();
    if (device < 0 || device >= ndev)
        return fail(PP_ERR_INVALID, "pp_create: no HIP device " + std::to_string(device) +
                                        " (" + std::to_string(ndev) + " visible)");
    pp_config c{};
    if (cfg) c = *cfg;
    if (int rc = normalise_config(c, "pp_create")) return rc;
    auto* ctx = new pp_ctx();
    ctx->device = device;
    ctx->cfg = c;
    auto cleanup = [&](int code) {
        pp_destroy(ctx);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return cleanup(fail(PP_ERR_HIP, "hipSetDevice failed"));
    if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
        hipSuccess)
        return cleanup(fail(PP_ERR_HIP, "hipDeviceGetAttribute failed"));
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess)
        return cleanup(fail(PP_ERR_HIP, "stream/event creation failed"));
    const size_t cap = (size_t)hit_capacity;
    const size_t bytes = cap * sizeof(HitRecord);
    if (hipMalloc(&ctx->hit_mem, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return cleanup(fail(PP_ERR_NOMEM, "pp_create: cannot allocate " + std::to_string(bytes) +
                                              " bytes of hit buffer"));
    }
    ctx->hits.rec = (HitRecord*)ctx->hit_mem;
    ctx->hits.capacity = cap;
    if (hipMalloc(&ctx->d_counters, sizeof(Counters)) != hipSuccess)
        return cleanup(fail(PP_ERR_NOMEM, "pp_create: counters"));
    *out = ctx;
    return PP_OK;
}

int pp_destroy(pp_ctx* ctx) {
    if (!ctx) return PP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->hit_mem) (void)hipFree(ctx->hit_mem);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_spec) (void)hipFree(ctx->d_spec);
    if (ctx->d_targets) (void)hipFree(ctx->d_targets);
    if (ctx->d_cells) (void)hipFree(ctx->d_cells);
    if (ctx->d_cell_items) (void)hipFree(ctx->d_cell_items);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    if (ctx->d_batch) (void)hipFree(ctx->d_batch);
    if (ctx->h_batch) (void)hipHostFree(ctx->h_batch);
    if (ctx->h_rec) (void)hipHostFree(ctx->h_rec);
    if (ctx->d_sort) (void)hipFree(ctx->d_sort);
    if (ctx->d_tgt_r) (void)hipFree(ctx->d_tgt_r);
    if (ctx->d_pmt) (void)hipFree(ctx->d_pmt);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PP_OK;
}

int pp_set_medium(pp_ctx* ctx, const pp_medium* m) {
    if (!ctx || !m) return fail(PP_ERR_INVALID, "pp_set_medium: NULL argument");
    if (m->kind != PP_MEDIUM_SIMPLE && m->kind != PP_MEDIUM_CASCADIA && m->kind != PP_MEDIUM_ICE)
        return fail(PP_ERR_INVALID, "pp_set_medium: unknown medium kind");
    if (m->phase_kind != PP_PHASE_HG && m->phase_kind != PP_PHASE_MIXED_HG_ES &&
        m->phase_kind != PP_PHASE_MIXED_HG_SL)
        return fail(PP_ERR_INVALID, "pp_set_medium: unknown phase function kind");
    if (m->phase_kind == PP_PHASE_MIXED_HG_SL && !(m->hg_fraction >= 0.0f && m->hg_fraction <= 1.0f))
        return fail(PP_ERR_INVALID, "pp_set_medium: hg_fraction in [0,1] required");
    if (m->kind == PP_MEDIUM_ICE &&
        !(m->ice_b_dust_400 > 0.0f && std::isfinite(m->ice_alpha_sca_dust) && m->sca_scale > 0.0f))
        return fail(PP_ERR_INVALID, "pp_set_medium: ICE needs b_dust_400 > 0, sca_scale > 0");
    if (!(m->hg_g > 0.0f && m->hg_g < 1.0f))
        return fail(PP_ERR_INVALID, "pp_set_medium: hg_g must be in (0,1)");
    if (m->phase_kind == PP_PHASE_MIXED_HG_ES &&
        !(m->es_b > 0.0f && m->hg_fraction >= 0.0f && m->hg_fraction <= 1.0f))
        return fail(PP_ERR_INVALID, "pp_set_medium: es_b > 0 and hg_fraction in [0,1] required");
    if (m->kind == PP_MEDIUM_SIMPLE &&
        !(m->sca_len > 0.0f && m->group_velocity > 0.0f && m->phase_refractive_index >= 1.0f))
        return fail(PP_ERR_INVALID, "pp_set_medium: SIMPLE needs sca_len>0, group_velocity>0, n>=1");
    if (m->kind == PP_MEDIUM_CASCADIA && !(m->sca_scale > 0.0f))
        return fail(PP_ERR_INVALID, "pp_set_medium: sca_scale must be > 0");
    ctx->medium_in = *m;
    ctx->medium = lower_medium(*m);
    ctx->has_medium = true;
    return PP_OK;
}

int pp_set_spectrum(pp_ctx* ctx, float mono_wl, const float* table, int n, int mapping) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_set_spectrum: NULL ctx");
    if (int rc = set_dev(ctx)) return rc;
    if (!table) {
        if (!(mono_wl > 0.0f) || !std::isfinite(mono_wl))
            return fail(PP_ERR_INVALID, "pp_set_spectrum: monochromatic wavelength must be > 0");
        ctx->spec_mono = mono_wl;
        ctx->spec_n = 0;
        ctx->has_spectrum = true;
        return PP_OK;
    }
    if (n < 2 || n > 8192) return fail(PP_ERR_INVALID, "pp_set_spectrum: n must be in [2, 8192]");
    if (mapping != PP_SPEC_KNOTS && mapping != PP_SPEC_TEXTURE)
        return fail(PP_ERR_INVALID, "pp_set_spectrum: unknown mapping");
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(table[i]) || table[i] <= 0.0f)
            return fail(PP_ERR_INVALID, "pp_set_spectrum: knots must be finite and > 0");
    if (n > ctx->spec_cap) {
        if (ctx->d_spec) PP_HIP(hipFree(ctx->d_spec));
        ctx->d_spec = nullptr;
        PP_HIP(hipMalloc(&ctx->d_spec, (size_t)n * sizeof(float)));
        ctx->spec_cap = n;
    }
    PP_HIP(hipMemcpyAsync(ctx->d_spec, table, (size_t)n * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    ctx->spec_n = n;
    ctx->spec_map = mapping;
    ctx->has_spectrum = true;
    return PP_OK;
}

int pp_set_targets(pp_ctx* ctx, const pp_target* t, int n) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_set_targets: NULL ctx");
    if (n < 0 || (n > 0 && !t)) return fail(PP_ERR_INVALID, "pp_set_targets: bad target list");
    bool planar = false;
    for (int i = 0; i < n; ++i) {
        const pp_target& x = t[i];
        const std::string who = "pp_set_targets: target " + std::to_string(i);
        if (!finite3(x.position)) return fail(PP_ERR_INVALID, who + " has a non-finite position");
        if (x.shape_kind == PP_SHAPE_SPHERE || x.shape_kind == PP_SHAPE_DISC) {
            if (!(x.radius > 0.0f) || !std::isfinite(x.radius))
                return fail(PP_ERR_INVALID, who + " needs a finite radius > 0");
        } else if (x.shape_kind == PP_SHAPE_RECTANGLE) {
            if (!(x.length[0] > 0.0f) || !(x.length[1] > 0.0f) || !std::isfinite(x.length[0]) ||
                !std::isfinite(x.length[1]))
                return fail(PP_ERR_INVALID, who + " needs finite side lengths > 0");
        } else {
            return fail(PP_ERR_INVALID, who + " has an unknown shape_kind");
        }
        planar |= x.shape_kind != PP_SHAPE_SPHERE;
    }
    // records per target: (x, y, z, r^2), plus (kind bits, lx/2, ly/2, 0) in a
    // list with planar targets
    const int per = planar ? 2 : 1;
    // planar lists are scanned from LDS; sphere lists of any length up to
    // kMaxTargets are searched by the grid, from device memory if need be
    if (planar && (size_t)n * per * sizeof(float4) + 8192 * sizeof(float) > (size_t)kMaxLdsBytes)
        return fail(PP_ERR_INVALID, "pp_set_targets: too many planar targets for LDS staging");
    if (n > kMaxTargets) return fail(PP_ERR_INVALID, "pp_set_targets: more than 65535 targets");
    if (int rc = set_dev(ctx)) return rc;
    std::vector<float4> h((size_t)std::max(n * per, 1));
    for (int i = 0; i < n; ++i) {
        const pp_target& x = t[i];
        // r^2 and the half lengths are Float32 values, as target_shape.radius^2
        // and length_x/2 of the reference
        const float w = x.shape_kind == PP_SHAPE_RECTANGLE ? 0.0f : x.radius * x.radius;
        h[(size_t)i * per] = make_float4(x.position[0], x.position[1], x.position[2], w);
        if (planar) {
            uint32_t kind = (uint32_t)x.shape_kind;
            float kf;
            std::memcpy(&kf, &kind, 4);
            h[(size_t)i * per + 1] = x.shape_kind == PP_SHAPE_RECTANGLE
                                         ? make_float4(kf, x.length[0] / 2.0f, x.length[1] / 2.0f, 0.0f)
                                         : make_float4(kf, 0.0f, 0.0f, 0.0f);
        }
    }
    // (x, y, z, radius) per target for the PMT detection of hits
    std::vector<float4> hr((size_t)std::max(n, 1));
    for (int i = 0; i < n; ++i)
        hr[(size_t)i] = make_float4(t[i].position[0], t[i].position[1], t[i].position[2],
                                    t[i].radius);
    if (n * per > ctx->tgt_cap) {
        if (ctx->d_targets) PP_HIP(hipFree(ctx->d_targets));
        if (ctx->d_tgt_r) PP_HIP(hipFree(ctx->d_tgt_r));
        ctx->d_targets = ctx->d_tgt_r = nullptr;
        ctx->tgt_cap = 0;
        PP_HIP(hipMalloc(&ctx->d_targets, (size_t)n * per * sizeof(float4)));
        PP_HIP(hipMalloc(&ctx->d_tgt_r, (size_t)n * per * sizeof(float4)));
        ctx->tgt_cap = n * per;
    }
    if (n > 0) {
        PP_HIP(hipMemcpyAsync(ctx->d_targets, h.data(), (size_t)n * per * sizeof(float4),
                              hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(ctx->d_tgt_r, hr.data(), (size_t)n * sizeof(float4),
                              hipMemcpyHostToDevice, ctx->stream));
    }
    PP_HIP(hipStreamSynchronize(ctx->stream));
    ctx->n_targets = n;
    ctx->planar = planar;
    ctx->h_targets.resize((size_t)n);
    ctx->h_radius.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        ctx->h_targets[(size_t)i] = h[(size_t)i * per];
        ctx->h_radius[(size_t)i] = t[i].radius;
    }
    ctx->grid_cell_built = -1.0f;  // grid is rebuilt on the next propagation
    return PP_OK;
}

int pp_propagate(pp_ctx* ctx, const pp_source* sources, int n_sources, uint64_t seed,
                 int32_t n_steps, float photon_scaling, int64_t* n_hits_out,
                 int64_t* n_ph_sim_out) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_propagate: NULL ctx");
    if (n_sources < 0 || (n_sources > 0 && !sources))
        return fail(PP_ERR_INVALID, "pp_propagate: bad source list");
    if (!ctx->has_medium || !ctx->has_spectrum || ctx->n_targets < 0)
        return fail(PP_ERR_STATE, "pp_propagate: set medium, spectrum and targets first");
    if (ctx->n_targets > kMaxTargets)
        return fail(PP_ERR_INVALID, "pp_propagate: more than 65535 targets");
    if (n_steps < 0 || n_steps > PP_MAX_STEPS)
        return fail(PP_ERR_INVALID, "pp_propagate: n_steps out of range");
    if (!(photon_scaling >= 0.0f) || !std::isfinite(photon_scaling))
        return fail(PP_ERR_INVALID, "pp_propagate: photon_scaling must be finite and >= 0");
    for (int i = 0; i < n_sources; ++i) {
        const pp_source& s = sources[i];
        if (s.kind < 0 || s.kind >= PP_SRC_KIND_COUNT)
            return fail(PP_ERR_INVALID, "pp_propagate: source " + std::to_string(i) +
                                            " has an unknown kind");
        if (s.photons < 0)
            return fail(PP_ERR_INVALID, "pp_propagate: negative photon count");
        if (!finite3(s.position) || !finite3(s.direction))
            return fail(PP_ERR_INVALID, "pp_propagate: non-finite source position/direction");
        const bool cascade = s.kind == PP_SRC_EXTENDED_CHERENKOV ||
                             s.kind == PP_SRC_EXTENDED_CHERENKOV_NO_SMEAR ||
                             s.kind == PP_SRC_EXTENDED_CHERENKOV_CUSTOM_SMEAR;
        if (cascade && !(s.long_a > 0.0f && s.long_b > 0.0f && s.long_lrad > 0.0f))
            return fail(PP_ERR_INVALID, "pp_propagate: cascade needs long_a, long_b, lrad > 0");
        if (s.kind == PP_SRC_EXTENDED_CHERENKOV_CUSTOM_SMEAR &&
            !(s.beta_a > 0.0f && s.beta_b > 0.0f))
            return fail(PP_ERR_INVALID, "pp_propagate: custom smear needs beta_a, beta_b > 0");
    }
    if (int rc = set_dev(ctx)) return rc;

    // search: grid for more than a handful of targets unless asked otherwise
    int search = ctx->cfg.search;
    if (ctx->planar && search == PP_SEARCH_GRID)
        return fail(PP_ERR_INVALID,
                    "pp_propagate: the grid search takes spherical targets only (the list holds "
                    "rectangles or discs)");
    const bool automatic = search == PP_SEARCH_AUTO;
    if (automatic)
        search = ctx->n_targets > 8 && !ctx->planar ? PP_SEARCH_GRID : PP_SEARCH_BRUTE;
    if (ctx->n_targets == 0) search = PP_SEARCH_BRUTE;
    if (search == PP_SEARCH_GRID && n_steps > kDeferredMaxSteps) {
        if (!automatic)
            return fail(PP_ERR_INVALID, "pp_propagate: the grid search takes at most 2^26 - 1 steps");
        search = PP_SEARCH_BRUTE;
    }
    if (search == PP_SEARCH_GRID) {
        const float want = ctx->cfg.grid_cell, want_z = ctx->cfg.grid_cell_z;
        if (ctx->grid_cell_built != want || ctx->grid_cell_z_built != want_z ||
            ctx->grid_cell_built < -0.5f) {
            if (int rc = build_grid(ctx, want, want_z)) {
                if (!automatic) return rc;
                search = PP_SEARCH_BRUTE;  // no grid fits: the every-target scan
            }
        }
    }
    const bool grid = search == PP_SEARCH_GRID;
    const bool global = grid && ctx->grid_global;
    const int skind = global ? kSearchGridGlobal
                      : grid ? kSearchGrid : ctx->planar ? kSearchMixed : kSearchBrute;
    // the global grid stages only the spectrum
    size_t lds = global ? std::max<size_t>((size_t)ctx->spec_n * 4, 16)
                        : lds_bytes(ctx->planar ? 2 * ctx->n_targets : ctx->n_targets,
                                    grid ? ctx->grid.n_cells : 0,
                                    grid ? ctx->grid.n_items : 0, ctx->spec_n, grid);
    // the grid kernels' per-wave walk queues follow the tables
    const uint32_t queue_off = (uint32_t)((lds + 15) & ~(size_t)15);
    // block size: the set one or, for 0, the one of 256, 384, 512, 768 and 1024
    // threads with the most waves resident per CU: the runtime's occupancy
    // query (registers: 6 waves per SIMD) and the LDS a workgroup needs (the
    // tables once and, with a grid, a walk queue per wave) in 512-byte
    // allocation units, which the query does not apply (it reports three
    // workgroups of 54320 bytes, and they run as two); the smallest on a tie.
    // (1000 targets: 768 threads, two workgroups of 12 waves, where 256 and
    // 512 threads leave 16 waves.)
    int block = ctx->cfg.block_size;
    if (block == 0) {
        const void* fn = batch_kernel_fn(skind, ctx->medium.phase.kind);
        int best_waves = 0;
        for (int b : {256, 384, 512, 768, 1024}) {
            const size_t need = grid ? queue_off + (size_t)(b / 64) * kWaveQueueBytes : lds;
            if (need > (size_t)kMaxLdsBytes) continue;
            int occ = 0;
            PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, b, need));
            const size_t alloc = (need + 511) & ~(size_t)511;
            occ = std::min<int>(occ, (int)((size_t)kMaxLdsBytes / alloc));
            if (occ * (b / 64) > best_waves) {
                best_waves = occ * (b / 64);
                block = b;
            }
        }
        if (block == 0) block = 64;  // the LDS check below reports what does not fit
    }
    ctx->block = block;
    if (grid) lds = queue_off + (size_t)(block / 64) * kWaveQueueBytes;
    if (lds > (size_t)kMaxLdsBytes)
        return fail(PP_ERR_INVALID, "pp_propagate: targets, tables and walk queues exceed the LDS "
                                    "of a workgroup (a smaller block_size may fit)");
    const int phase = ctx->medium.phase.kind;
    const bool count = ctx->cfg.count_work != 0;
    ctx->last_search = search;
    KParams kp{};
    kp.med = ctx->medium;
    kp.k0 = (uint32_t)seed;
    kp.k1 = (uint32_t)(seed >> 32);
    kp.n_steps = n_steps;
    kp.queue_off = queue_off;
    kp.targets = ctx->d_targets;
    kp.n_targets = ctx->n_targets;
    if (grid) {
        kp.grid = ctx->grid;
        kp.cells = ctx->d_cells;
        kp.cell_items = ctx->d_cell_items;
    }
    kp.spec_n = ctx->spec_n;
    kp.spec_map = ctx->spec_map;
    kp.spec_mono = ctx->spec_mono;
    kp.spec_nf = (float)ctx->spec_n;
    kp.spec_lastf = (float)(ctx->spec_n - 1);
    kp.spec = ctx->d_spec;
    kp.hits = ctx->hits;
    kp.counters = ctx->d_counters;
    // photon count of src/photon_prop_cuda.jl:456 (Float32 product)
    std::vector<uint64_t> nph((size_t)n_sources);
    for (int i = 0; i < n_sources; ++i) {
        // the limit is checked on the float product, before it is converted
        const float prod = std::floor((float)sources[i].photons * photon_scaling);
        if (!(prod < (float)kMaxPhotonsPerSource))
            return fail(PP_ERR_INVALID, "pp_propagate: source " + std::to_string(i) +
                                            " has 2^40 or more photons after scaling");
        nph[(size_t)i] = prod > 0.0f ? (uint64_t)prod : 0;
    }
    ctx->n_launch = 0;
    const int mode = ctx->cfg.batch_mode == PP_BATCH_AUTO ? PP_BATCH_PERSISTENT : ctx->cfg.batch_mode;
    if (mode == PP_BATCH_PER_SOURCE) {
        if (int rc = launch_per_source(ctx, kp, sources, nph, skind, phase, count, lds)) return rc;
    } else {
        if (int rc = launch_batch(ctx, kp, sources, nph, skind, phase, count, lds, mode == PP_BATCH_GRAPH))
            return rc;
    }
    // the grid kernels leave their hit records to be completed
    if (skind == kSearchGrid || skind == kSearchGridGlobal) {
        hipLaunchKernelGGL(replay_hits_kernel, dim3((unsigned)(4 * ctx->n_cu)), dim3(256), 0,
                           ctx->stream, kp, (const pp_source*)ctx->d_batch);
        PP_HIP(hipGetLastError());
    }
    PP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    Counters hc{};
    PP_HIP(hipMemcpyAsync(&hc, ctx->d_counters, sizeof hc, hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    PP_HIP(hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1));
    ctx->n_hits = (int64_t)std::min<unsigned long long>(hc.n_hits, ctx->hits.capacity);
    ctx->n_ph_sim = (int64_t)hc.n_ph_sim;
    ctx->work_steps = (int64_t)hc.steps;
    ctx->work_tests = (int64_t)hc.tests;
    ctx->work_cells = (int64_t)hc.cells;
    ctx->work_walks = (int64_t)hc.walks;
    ctx->perm_valid = false;
    ctx->sorted_valid = false;
    ctx->last_n_sources = n_sources;
    if (n_hits_out) *n_hits_out = ctx->n_hits;
    if (n_ph_sim_out) *n_ph_sim_out = ctx->n_ph_sim;
    if (hc.err) return fail(PP_ERR_HIT_OVERFLOW, "hit buffer overflow (PhotonPropOOMException)");
    return PP_OK;
}

int pp_get_config(pp_ctx* ctx, pp_config* cfg) {
    if (!ctx || !cfg) return fail(PP_ERR_INVALID, "pp_get_config: NULL argument");
    *cfg = ctx->cfg;
    return PP_OK;
}

int pp_set_config(pp_ctx* ctx, const pp_config* cfg) {
    if (!ctx || !cfg) return fail(PP_ERR_INVALID, "pp_set_config: NULL argument");
    pp_config c = *cfg;
    if (int rc = normalise_config(c, "pp_set_config")) return rc;
    ctx->cfg = c;
    return PP_OK;
}

int pp_autotune(pp_ctx* ctx, const pp_source* sources, int n_sources, uint64_t seed,
                int32_t n_steps, float photon_scaling, int64_t sample_photons, pp_config* best,
                float* best_ms) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_autotune: NULL ctx");
    if (n_sources <= 0 || !sources) return fail(PP_ERR_INVALID, "pp_autotune: no sources");
    if (sample_photons <= 0) return fail(PP_ERR_INVALID, "pp_autotune: sample_photons must be > 0");
    if (!(photon_scaling > 0.0f) || !std::isfinite(photon_scaling))
        return fail(PP_ERR_INVALID, "pp_autotune: photon_scaling must be finite and > 0");
    double total = 0.0;
    for (int i = 0; i < n_sources; ++i) total += (double)sources[i].photons * photon_scaling;
    if (!(total > 0.0)) return fail(PP_ERR_INVALID, "pp_autotune: the sources hold no photons");
    const float scale = (float)std::min<double>(photon_scaling,
                                                photon_scaling * (double)sample_photons / total);
    const pp_config saved = ctx->cfg;
    pp_config cur = saved;
    if (cur.batch_mode == PP_BATCH_AUTO) cur.batch_mode = PP_BATCH_PERSISTENT;
    cur.count_work = 0;
    cur.waves_per_simd = 0;
    auto run = [&](const pp_config& c, float& ms) -> int {
        ctx->cfg = c;
        float t = INFINITY;
        for (int r = 0; r < 3; ++r) {
            int64_t nh = 0, ns = 0;
            const int rc = pp_propagate(ctx, sources, n_sources, seed, n_steps, scale, &nh, &ns);
            if (rc != PP_OK) return rc;
            if (r > 0) t = std::min(t, ctx->last_ms);
        }
        ms = t;
        return PP_OK;
    };
    float t_best = 0.0f;
    if (int rc = run(cur, t_best)) {
        ctx->cfg = saved;
        return rc;
    }
    // automatic grid edge of these targets, for the cell candidates
    const bool grid = ctx->last_search == PP_SEARCH_GRID;
    const float h_auto = grid ? ctx->grid.h : 0.0f;
    std::vector<float> cells = {0.0f};
    if (grid)
        for (float f : {0.6f, 0.8f, 0.9f, 1.1f, 1.25f}) cells.push_back(f * h_auto);
    const int blocks[] = {64, 128, 256, 384, 512, 768, 1024};
    const int chunks[] = {0, 64, 128, 256, 512, 1024};
    // cell shapes: cubic, columns, and cells twice as tall as wide (set when tried)
    float shapes[] = {0.0f, -1.0f, 0.0f};
    for (int pass = 0; pass < 2; ++pass) {
        for (int knob = 0; knob < 4; ++knob) {
            if (knob >= 2 && !grid) continue;
            const size_t n_cand = knob == 0 ? 7 : knob == 1 ? 6 : knob == 2 ? cells.size() : 3;
            if (knob == 3) shapes[2] = 2.0f * (cur.grid_cell > 0.0f ? cur.grid_cell : h_auto);
            for (size_t k = 0; k < n_cand; ++k) {
                pp_config c = cur;
                if (knob == 0) c.block_size = blocks[k];
                if (knob == 1) c.chunk_photons = chunks[k];
                if (knob == 2) c.grid_cell = cells[k];
                if (knob == 3) c.grid_cell_z = shapes[k];
                if (c.block_size == cur.block_size && c.chunk_photons == cur.chunk_photons &&
                    c.grid_cell == cur.grid_cell && c.grid_cell_z == cur.grid_cell_z)
                    continue;
                if (c.batch_mode == PP_BATCH_PER_SOURCE && knob == 1) continue;
                float t = 0.0f;
                const int rc = run(c, t);
                if (rc == PP_ERR_HIT_OVERFLOW || rc == PP_ERR_INVALID) continue;  // e.g. LDS budget
                if (rc != PP_OK) {
                    ctx->cfg = saved;
                    return rc;
                }
                if (t < t_best) {
                    t_best = t;
                    cur = c;
                }
            }
        }
    }
    cur.count_work = saved.count_work;
    cur.deterministic_order = saved.deterministic_order;
    ctx->cfg = cur;
    ctx->n_hits = 0;  // the tuning runs' hits are not results
    ctx->n_ph_sim = 0;
    ctx->perm_valid = false;
    ctx->sorted_valid = false;
    if (best) *best = cur;
    if (best_ms) *best_ms = t_best;
    return PP_OK;
}

int pp_last_timing(pp_ctx* ctx, float* kernel_ms, int32_t* n_launches) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_last_timing: NULL ctx");
    if (kernel_ms) *kernel_ms = ctx->last_ms;
    if (n_launches) *n_launches = ctx->n_launch;
    return PP_OK;
}

void* pp_stream(pp_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int pp_work_counts(pp_ctx* ctx, int64_t* steps, int64_t* sphere_tests, int64_t* cell_visits,
                   int64_t* grid_walks) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_work_counts: NULL ctx");
    if (!ctx->cfg.count_work)
        return fail(PP_ERR_STATE, "pp_work_counts: context created without count_work");
    if (steps) *steps = ctx->work_steps;
    if (sphere_tests) *sphere_tests = ctx->work_tests;
    if (cell_visits) *cell_visits = ctx->work_cells;
    if (grid_walks) *grid_walks = ctx->work_walks;
    return PP_OK;
}

int pp_search_info(pp_ctx* ctx, int32_t* search, float* cell, int32_t* dims, int32_t* n_items) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_search_info: NULL ctx");
    if (search) *search = ctx->last_search;
    const bool g = ctx->last_search == PP_SEARCH_GRID && ctx->grid_cell_built > -0.5f;
    if (cell) *cell = g ? ctx->grid.h : 0.0f;
    if (dims) {
        dims[0] = g ? ctx->grid.nx : 0;
        dims[1] = g ? ctx->grid.ny : 0;
        dims[2] = g ? ctx->grid.nz : 0;
    }
    if (n_items) *n_items = g ? ctx->grid.n_items : 0;
    return PP_OK;
}

// Sort key of hit k: the source index above the 40 photon-id bits, so that an
// ascending sort orders by source, then by photon id; the value is k.
__global__ void hit_keys_kernel(const HitRecord* __restrict__ rec, int64_t n,
                                uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    key[k] = ((uint64_t)(uint32_t)rec[k].source_ix << kPidBits) |
             (rec[k].id & (kMaxPhotonsPerSource - 1));
    idx[k] = (uint32_t)k;
}

// out[k] = rec[idx[k]], 64 bytes per lane as four 16-byte loads and stores.
__global__ void gather_hits_kernel(const HitRecord* __restrict__ rec,
                                   const uint32_t* __restrict__ idx, int64_t n,
                                   HitRecord* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint4* src = (const uint4*)(rec + idx[k]);
    uint4* dst = (uint4*)(out + k);
    const uint4 a = src[0], b = src[1], c = src[2], d = src[3];
    dst[0] = a;
    dst[1] = b;
    dst[2] = c;
    dst[3] = d;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The hits in (source, photon id) order, written by the device to the front
// of ctx->d_sort: a radix sort over 40 + bits(n_sources - 1) key bits carries
// the record indices (the sort is stable; a photon hits at most once, so keys
// are unique), then a gather writes the records in that order.  Sets
// *done = false and does nothing when the keys do not fit 64 bits, the
// indices 32 bits or the buffer device memory; the caller then sorts on the
// host.
static int sort_hits_device(pp_ctx* ctx, bool* done) {
    *done = false;
    const int64_t n = ctx->n_hits;
    const int32_t ns = ctx->last_n_sources;
    if (n <= 0 || n > INT32_MAX || ns <= 0 || ns > (1 << 24)) return PP_OK;
    int end_bit = (int)kPidBits;
    for (uint32_t v = (uint32_t)(ns - 1); v; v >>= 1) ++end_bit;
    size_t temp = 0;
    PP_HIP(rocprim::radix_sort_pairs(nullptr, temp, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t)n, 0u,
                                     (unsigned)end_bit, ctx->stream));
    const size_t br = align256((size_t)n * sizeof(HitRecord));
    const size_t bk = align256((size_t)n * 8), bv = align256((size_t)n * 4);
    const size_t need = br + 2 * bk + 2 * bv + align256(temp);
    if (need > ctx->d_sort_cap) {
        if (ctx->d_sort) PP_HIP(hipFree(ctx->d_sort));
        ctx->d_sort = nullptr;
        ctx->d_sort_cap = 0;
        if (hipMalloc((void**)&ctx->d_sort, need) != hipSuccess) {
            (void)hipGetLastError();  // out of device memory: sort on the host
            ctx->d_sort = nullptr;
            return PP_OK;
        }
        ctx->d_sort_cap = need;
    }
    char* base = ctx->d_sort + br;
    uint64_t* k_in = (uint64_t*)base;
    uint64_t* k_out = (uint64_t*)(base + bk);
    uint32_t* v_in = (uint32_t*)(base + 2 * bk);
    uint32_t* v_out = (uint32_t*)(base + 2 * bk + bv);
    void* tmp = base + 2 * bk + 2 * bv;
    hit_keys_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(ctx->hits.rec, n, k_in, v_in);
    PP_HIP(hipGetLastError());
    PP_HIP(rocprim::radix_sort_pairs(tmp, temp, k_in, k_out, v_in, v_out, (uint32_t)n, 0u,
                                     (unsigned)end_bit, ctx->stream));
    gather_hits_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
        ctx->hits.rec, v_out, n, (HitRecord*)ctx->d_sort);
    PP_HIP(hipGetLastError());
    *done = true;
    return PP_OK;
}

int pp_copy_hits(pp_ctx* ctx, int64_t first, int64_t n, float* pos3, float* dir3,
                 float* init_dir3, float* wavelength, void* time, int time_f64,
                 float* dist_travelled, int64_t* module_ix, int32_t* n_steps,
                 uint64_t* photon_id, int32_t* source_ix) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_copy_hits: NULL ctx");
    if (first < 0 || n < 0 || first > ctx->n_hits || n > ctx->n_hits - first)
        return fail(PP_ERR_INVALID, "pp_copy_hits: range outside [0, n_hits)");
    if (n == 0) return PP_OK;
    if (int rc = set_dev(ctx)) return rc;
    // deterministic_order: 0 = buffer order, 1 = sorted on the device (host
    // where the device cannot), 2 = sorted on the host (no device buffer)
    const bool ordered = ctx->cfg.deterministic_order != 0;
    if (ordered && ctx->cfg.deterministic_order != 2 && !ctx->sorted_valid && !ctx->perm_valid) {
        bool done = false;
        if (int rc = sort_hits_device(ctx, &done)) return rc;
        ctx->sorted_valid = done;
    }
    // the host sorts only where the device could not: it then reads every record
    const bool host_sort = ordered && (ctx->cfg.deterministic_order == 2 || !ctx->sorted_valid);
    // records to read: [lo, lo+cnt) of the device buffer src
    const HitRecord* src = ordered && !host_sort ? (const HitRecord*)ctx->d_sort : ctx->hits.rec;
    const int64_t lo = host_sort ? 0 : first, cnt = host_sort ? ctx->n_hits : n;
    // records stream through a pinned staging buffer of at most
    // kStageRecords records (kept by the context); the host sort, which
    // needs random access to every record, collects them in pageable memory
    const int64_t stage = std::min<int64_t>(cnt, kStageRecords);
    if (stage > ctx->h_rec_cap) {
        if (ctx->h_rec) PP_HIP(hipHostFree(ctx->h_rec));
        ctx->h_rec = nullptr;
        ctx->h_rec_cap = 0;
        PP_HIP(hipHostMalloc((void**)&ctx->h_rec, (size_t)stage * sizeof(HitRecord),
                             hipHostMallocDefault));
        ctx->h_rec_cap = stage;
    }
    const uint64_t pid_mask = kMaxPhotonsPerSource - 1;
    auto scatter = [&](const HitRecord& r, int64_t k) {
        for (int c = 0; c < 3; ++c) {
            if (pos3) pos3[3 * k + c] = r.f[c];
            if (dir3) dir3[3 * k + c] = r.f[3 + c];
            if (init_dir3) init_dir3[3 * k + c] = r.f[6 + c];
        }
        if (wavelength) wavelength[k] = r.f[9];
        if (time) {
            if (time_f64)
                ((double*)time)[k] = r.f[10];
            else
                ((float*)time)[k] = r.f[10];
        }
        if (dist_travelled) dist_travelled[k] = r.f[11];
        if (module_ix) module_ix[k] = (int64_t)((r.id >> kPidBits) & 0xFFFF);
        if (n_steps) n_steps[k] = r.n_steps;
        if (photon_id) photon_id[k] = r.id & pid_mask;
        if (source_ix) source_ix[k] = r.source_ix;
    };
    std::vector<HitRecord> all;
    if (host_sort) all.resize((size_t)cnt);
    for (int64_t off = 0; off < cnt; off += stage) {
        const int64_t m = std::min<int64_t>(stage, cnt - off);
        PP_HIP(hipMemcpyAsync(ctx->h_rec, src + lo + off, (size_t)m * sizeof(HitRecord),
                              hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (host_sort)
            std::memcpy(all.data() + off, ctx->h_rec, (size_t)m * sizeof(HitRecord));
        else
            for (int64_t k = 0; k < m; ++k) scatter(ctx->h_rec[k], off + k);
    }
    if (!host_sort) return PP_OK;
    if (!ctx->perm_valid) {
        // sort (source, photon id) keys beside their record index: one
        // contiguous array instead of indirect two-field compares
        struct Key {
            int32_t src;
            uint64_t pid;
            int64_t idx;
        };
        std::vector<Key> keys((size_t)ctx->n_hits);
        for (int64_t k = 0; k < ctx->n_hits; ++k)
            keys[(size_t)k] = {all[(size_t)k].source_ix, all[(size_t)k].id & pid_mask, k};
        std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
            return a.src != b.src ? a.src < b.src : a.pid < b.pid;
        });
        ctx->perm.resize((size_t)ctx->n_hits);
        for (int64_t k = 0; k < ctx->n_hits; ++k) ctx->perm[(size_t)k] = keys[(size_t)k].idx;
        ctx->perm_valid = true;
    }
    for (int64_t k = 0; k < n; ++k) scatter(all[(size_t)ctx->perm[(size_t)(first + k)]], k);
    return PP_OK;
}

}  // extern "C"

int pp::ctx_hit_records(pp_ctx* ctx, const HitRecord** rec, int64_t* n, hipStream_t* stream,
                        int* device) {
    if (!ctx) return fail(PP_ERR_INVALID, "NULL ctx");
    if (int rc = set_dev(ctx)) return rc;
    *rec = ctx->hits.rec;
    *n = ctx->n_hits;
    *stream = ctx->stream;
    *device = ctx->device;
    return PP_OK;
}

extern "C" {

int pp_release_scratch(pp_ctx* ctx) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_release_scratch: NULL ctx");
    if (int rc = set_dev(ctx)) return rc;
    PP_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->d_sort) PP_HIP(hipFree(ctx->d_sort));
    if (ctx->h_rec) PP_HIP(hipHostFree(ctx->h_rec));
    if (ctx->d_pmt) PP_HIP(hipFree(ctx->d_pmt));
    ctx->d_sort = nullptr;
    ctx->d_sort_cap = 0;
    ctx->sorted_valid = false;
    ctx->h_rec = nullptr;
    ctx->h_rec_cap = 0;
    ctx->d_pmt = nullptr;
    ctx->pmt_cap = 0;
    std::vector<int64_t>().swap(ctx->perm);
    ctx->perm_valid = false;
    return PP_OK;
}

int pp_pom_pmt_hits_ctx(pp_ctx* ctx, const pp_pom_model* m, const float* wl_grid,
                        const float* acc1, const float* acc2, uint64_t seed, int64_t first,
                        int64_t n, int32_t* pmt_id_out, int64_t* n_detected, float* kernel_ms) {
    if (!ctx) return fail(PP_ERR_INVALID, "pp_pom_pmt_hits_ctx: NULL ctx");
    if (n_detected) *n_detected = 0;
    if (kernel_ms) *kernel_ms = 0.0f;
    if (!m || m->n_wl < 2 || m->n_wl > 1024)
        return fail(PP_ERR_INVALID, "pp_pom_pmt_hits_ctx: model with n_wl in [2, 1024] required");
    if (first < 0 || n < 0 || first > ctx->n_hits || n > ctx->n_hits - first)
        return fail(PP_ERR_INVALID, "pp_pom_pmt_hits_ctx: range outside [0, n_hits)");
    if (ctx->planar)
        return fail(PP_ERR_INVALID,
                    "pp_pom_pmt_hits_ctx: the target list holds planar targets (POMs are spheres)");
    pp::PmtSetup st{};
    std::vector<float> tab(3 * (size_t)m->n_wl);
    if (int rc = pp::pmt_setup(m, wl_grid, acc1, acc2, seed, &st, tab.data())) return rc;
    if (n == 0) return PP_OK;
    if (!pmt_id_out) return fail(PP_ERR_INVALID, "pp_pom_pmt_hits_ctx: NULL output");
    if (int rc = set_dev(ctx)) return rc;
    // the records in the order pp_copy_hits returns them
    const bool ordered = ctx->cfg.deterministic_order != 0;
    if (ordered && ctx->cfg.deterministic_order != 2 && !ctx->sorted_valid && !ctx->perm_valid) {
        bool done = false;
        if (int rc = sort_hits_device(ctx, &done)) return rc;
        ctx->sorted_valid = done;
    }
    const bool host_sort = ordered && (ctx->cfg.deterministic_order == 2 || !ctx->sorted_valid);
    const HitRecord* src = ordered && !host_sort ? (const HitRecord*)ctx->d_sort : ctx->hits.rec;
    const int64_t lo = host_sort ? 0 : first, cnt = host_sort ? ctx->n_hits : n;
    // table | detected count | PMT ids, kept by the context and grown when needed
    const size_t off_cnt = align256(tab.size() * sizeof(float));
    const size_t off_out = off_cnt + 256;
    const size_t need = off_out + (size_t)cnt * sizeof(int32_t);
    if (need > ctx->pmt_cap) {
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->d_pmt) PP_HIP(hipFree(ctx->d_pmt));
        ctx->d_pmt = nullptr;
        ctx->pmt_cap = 0;
        if (hipMalloc((void**)&ctx->d_pmt, need) != hipSuccess) {
            (void)hipGetLastError();
            return fail(PP_ERR_NOMEM, "pp_pom_pmt_hits_ctx: device allocation failed");
        }
        ctx->pmt_cap = need;
    }
    float* d_tab = (float*)ctx->d_pmt;
    auto* d_cnt = (unsigned long long*)(ctx->d_pmt + off_cnt);
    auto* d_out = (int32_t*)(ctx->d_pmt + off_out);
    PP_HIP(hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    PP_HIP(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), ctx->stream));
    PP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (int rc = pp::pmt_launch_records(st, d_tab, src + lo, cnt, ctx->d_tgt_r, d_out, d_cnt,
                                        ctx->stream, ctx->n_cu))
        return rc;
    PP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<int32_t> all;
    int32_t* dst = pmt_id_out;
    if (host_sort) {
        all.resize((size_t)cnt);
        dst = all.data();
    }
    PP_HIP(hipMemcpyAsync(dst, d_out, (size_t)cnt * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    if (kernel_ms) PP_HIP(hipEventElapsedTime(kernel_ms, ctx->ev0, ctx->ev1));
    if (host_sort) {
        if (!ctx->perm_valid) {
            // the same (source, photon id) order as pp_copy_hits's host sort
            std::vector<HitRecord> rec((size_t)ctx->n_hits);
            PP_HIP(hipMemcpyAsync(rec.data(), ctx->hits.rec, rec.size() * sizeof(HitRecord),
                                  hipMemcpyDeviceToHost, ctx->stream));
            PP_HIP(hipStreamSynchronize(ctx->stream));
            std::vector<int64_t> perm((size_t)ctx->n_hits);
            std::iota(perm.begin(), perm.end(), 0);
            const uint64_t mask = kMaxPhotonsPerSource - 1;
            std::sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
                const HitRecord &x = rec[(size_t)a], &y = rec[(size_t)b];
                return x.source_ix != y.source_ix ? x.source_ix < y.source_ix
                                                  : (x.id & mask) < (y.id & mask);
            });
            ctx->perm.swap(perm);
            ctx->perm_valid = true;
        }
        for (int64_t k = 0; k < n; ++k) pmt_id_out[k] = all[(size_t)ctx->perm[(size_t)(first + k)]];
    }
    if (n_detected) {
        int64_t c = 0;
        for (int64_t k = 0; k < n; ++k) c += pmt_id_out[k] != 0;
        *n_detected = c;
    }
    return PP_OK;
}

int pp_eval_fn(pp_ctx* ctx, int fn, const float* in, int in_stride, float* out, int out_stride,
               int64_t n) {
    static const int kIn[PP_FN_COUNT] = {2, 6, 2, 2, 5, 11, 2, 1, 2, 6, 7, 1, 2, 13, 5};
    static const int kOut[PP_FN_COUNT] = {3, 3, 1, 1, 3, 2, 1, 3, 3, 4, 2, 1, 1, 2, 1};
    if (!ctx) return fail(PP_ERR_INVALID, "pp_eval_fn: NULL ctx");
    if (fn < 0 || fn >= PP_FN_COUNT) return fail(PP_ERR_INVALID, "pp_eval_fn: unknown function");
    if (n < 0 || (n > 0 && (!in || !out)))
        return fail(PP_ERR_INVALID, "pp_eval_fn: bad buffers");
    if (in_stride < kIn[fn] || out_stride < kOut[fn])
        return fail(PP_ERR_INVALID, "pp_eval_fn: stride smaller than the function's arity");
    if (fn == PP_FN_CASCADIA && !ctx->has_medium)
        return fail(PP_ERR_STATE, "pp_eval_fn: set a medium first");
    if (fn == PP_FN_SPECTRUM && !ctx->has_spectrum)
        return fail(PP_ERR_STATE, "pp_eval_fn: set a spectrum first");
    if (n == 0) return PP_OK;
    if (int rc = set_dev(ctx)) return rc;
    float *d_in = nullptr, *d_out = nullptr;
    const size_t in_b = (size_t)n * in_stride * 4, out_b = (size_t)n * out_stride * 4;
    PP_HIP(hipMalloc(&d_in, in_b));
    if (hipMalloc(&d_out, out_b) != hipSuccess) {
        (void)hipFree(d_in);
        return fail(PP_ERR_NOMEM, "pp_eval_fn: output allocation");
    }
    int rc = PP_OK;
    do {
        // everything on the context stream, in order
        if (hipMemcpyAsync(d_in, in, in_b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) { rc = fail(PP_ERR_HIP, "copy in"); break; }
        if (hipMemsetAsync(d_out, 0, out_b, ctx->stream) != hipSuccess) { rc = fail(PP_ERR_HIP, "memset"); break; }
        EvalParams e{};
        e.fn = fn;
        e.in_stride = in_stride;
        e.out_stride = out_stride;
        e.n = n;
        e.in = d_in;
        e.out = d_out;
        e.med = ctx->medium;
        e.spec = ctx->d_spec;
        e.spec_n = ctx->spec_n;
        e.spec_map = ctx->spec_map;
        e.spec_mono = ctx->spec_mono;
        e.spec_nf = (float)ctx->spec_n;
        e.spec_lastf = (float)(ctx->spec_n - 1);
        const int64_t grid = (n + 255) / 256;
        hipLaunchKernelGGL(eval_fn_kernel, dim3((unsigned)grid), dim3(256), 0, ctx->stream, e);
        if (hipGetLastError() != hipSuccess) { rc = fail(PP_ERR_HIP, "eval launch"); break; }
        if (hipMemcpyAsync(out, d_out, out_b, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) { rc = fail(PP_ERR_HIP, "copy out"); break; }
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) { rc = fail(PP_ERR_HIP, "eval sync"); break; }
    } while (0);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}

}  // extern "C"
