// persistent.hip -- nbl_driver_run_persistent: every primary of a pixel
// batch runs to the end of its cascade in ONE launch.
//
// Replaces the prescan + frame loop of gpu_main.cu:108-213 (push, sort,
// do_iteration x frame_size, flush) for throughput.  Design:
//   * each lane of a resident wavefront holds one electron in registers
//     (position, direction, energy, material, last triangle, pixel tag);
//     a lane's XORWOW stream persists across the electrons it handles;
//   * each wavefront keeps a LIFO stack of secondaries in LDS; a lane whose
//     electron ended pops a secondary first, otherwise the wavefront takes a
//     contiguous range of primaries with ONE atomic on a global counter
//     (compaction: idle lanes are refilled at the top of every step, or when
//     at least refill_threshold lanes are idle);
//   * an inelastic event may run only if the stack has room for the
//     secondary it may create; otherwise the electron waits (PENDING,
//     gpu_driver.inl:523-528) and retries next step without resampling;
//   * inelastic events run in batches: their code path (table draws,
//     energy-loss sampling, the secondary's direction) costs the whole
//     wavefront however few lanes take it, so a lane whose event is
//     inelastic waits in the same way until inel_batch lanes want one, one
//     has waited inel_wait iterations, or no live lane has other work;
//   * detections update per-pixel SE/BSE counters, a 1-eV histogram and the
//     work counters used for the roofline.
// Every wavefront exits once the primary counter is exhausted, its stack is
// empty and all its lanes are idle; each electron is capped at 2^20 steps.
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>

#include "driver.h"
#include "sampler.h"
#include "wave_propagate.h"

using namespace nbl;

namespace {

constexpr uint32_t STEP_CAP = 1u << 20;
constexpr int N_COUNTERS = 15;
// counter slots
enum { C_PRIM = 0, C_SE, C_BSE, C_EVENTS, C_L1D, C_L2D, C_RD, C_NODES, C_TRIS, C_BOUND, C_SEC, C_PEND, C_KILLED, C_DET,
       C_DEPN };

constexpr uint32_t N_QUEUES = 8;       // one primary queue per XCD (blocks go round-robin over XCDs)
constexpr uint32_t QSTRIDE = 16;       // queue counters 128 B apart

// A beam spot as the kernel sees it: a caller's pixel or one tile of it.
struct dpixel {
    float x, y, z, energy, sigma, half_x, half_y;
    uint32_t profile;
    uint32_t tally;                     // index of the caller's pixel
};

// One energy-deposition event (nbl_deposit_record).
struct drecord { f3 pos; float energy, deposit; };

struct pers_params {
    dgeom g;
    const dmat* mats;
    float min_e, max_e;
    const dpixel* pixels;
    const uint64_t* offsets;            // n_pixels + 1 prefix sums of n_primaries
    uint32_t n_pixels;
    uint64_t total;
    uint64_t qbeg[N_QUEUES + 1];        // primary index range of each queue
    f3 dir;
    uint64_t prim_seed, rng_seed, rng_first;
    unsigned long long* next;           // N_QUEUES counters, QSTRIDE apart
    uint32_t* pixel_se;
    uint32_t* pixel_bse;
    unsigned long long* hist;
    unsigned long long* counters;
    float* overflow;                    // per-wavefront HBM stack regions, 9 x ovf floats each
    uint32_t depth;                     // LDS stack entries per wavefront
    uint32_t ovf;                       // HBM stack entries per wavefront
    uint32_t refill_threshold;
    uint32_t inel_batch, inel_wait;     // inelastic batching (see above)
    dmlp mlp[2];                        // learned sampler (elastic, inelastic) when LEARNED
    uint32_t learned;                   // bit k: network k replaces table k (LEARNED)
    // explicit primaries (plist != nullptr): primary i is plist[i] with tag ptags[i]
    const dparticle* plist;
    const uint32_t* ptags;
    // detected-electron records (det_on == 0: off), in detection order
    uint32_t det_on;
    dparticle* det;
    uint32_t* det_tags;
    unsigned long long* det_count;      // all detections (may exceed det_cap)
    uint64_t det_cap;
    // energy-deposition tally (DEPOSIT): grid (dep == nullptr: off) and records (drec_on == 0: off)
    double* dep;                        // [dnz][dny][dnx] voxels
    double* dep_total;                  // all deposits (nullptr: not summed)
    uint32_t drec_on;
    drecord* drec;
    uint32_t* drec_tags;
    unsigned long long* drec_count;     // all deposit events (may exceed drec_cap)
    uint64_t drec_cap;
    uint32_t dnx, dny, dnz;
    f3 dlo, dinv;                       // voxel index = floor((p - dlo) * dinv)
};

// One logical stack per wavefront: entries [0, D) in LDS, [D, D + G) in HBM.
// Fields are stored SoA: x y z dx dy dz e mat tag.  The LDS part is typed as
// LDS (address space 3), so that its accesses are LDS instructions: through
// one generic pointer for both parts they would be flat accesses, which wait
// on the vector-memory and the LDS counters alike.
typedef __attribute__((address_space(3))) float lds_float;

struct wave_stack {
    lds_float* lds; float* hbm; uint32_t D, G;
    template<typename PTR>
    static __device__ __forceinline__ void store(PTR b, size_t s, const dparticle& q, int mat, uint32_t tag)
    {
        b[0] = q.pos.x; b[s] = q.pos.y; b[2 * s] = q.pos.z;
        b[3 * s] = q.dir.x; b[4 * s] = q.dir.y; b[5 * s] = q.dir.z;
        b[6 * s] = q.e; b[7 * s] = __int_as_float(mat); b[8 * s] = __uint_as_float(tag);
    }
    template<typename PTR>
    static __device__ __forceinline__ void load(PTR b, size_t s, dparticle& q, int& mat, uint32_t& tag)
    {
        q.pos = { b[0], b[s], b[2 * s] };
        q.dir = { b[3 * s], b[4 * s], b[5 * s] };
        q.e = b[6 * s]; mat = __float_as_int(b[7 * s]); tag = __float_as_uint(b[8 * s]);
    }
    __device__ __forceinline__ void put(uint32_t j, const dparticle& q, int mat, uint32_t tag) const
    {
        if (j < D) store(lds + j, D, q, mat, tag);
        else store(hbm + (j - D), G, q, mat, tag);
    }
    __device__ __forceinline__ void get(uint32_t j, dparticle& q, int& mat, uint32_t& tag) const
    {
        if (j < D) load(lds + j, D, q, mat, tag);
        else load(hbm + (j - D), G, q, mat, tag);
    }
};

__device__ __forceinline__ uint64_t bcast64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Primary i: pixel by binary search, position from a counter-based hash.
__device__ void make_primary(const pers_params& P, uint64_t i, dparticle& p, uint32_t& tag)
{
    if (P.plist) {                                      // explicit primaries
        p = P.plist[i];
        tag = P.ptags[i];
        return;
    }
    uint32_t lo = 0, hi = P.n_pixels;                    // offsets[lo] <= i < offsets[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (P.offsets[mid] <= i) lo = mid; else hi = mid;
    }
    const dpixel px = P.pixels[lo];
    const uint64_t h = mix64(P.prim_seed + 0x9E3779B97F4A7C15ull * (i + 1));
    const float u1 = (float)((uint32_t)h >> 8) * 5.9604644775390625e-8f + 2.98023223876953125e-8f;  // (0,1)
    const float u2 = (float)((uint32_t)(h >> 32) >> 8) * 5.9604644775390625e-8f;
    float x = px.x, y = px.y;
    if (px.profile == 0) {
        const float r = px.sigma * sqrtf(-2.0f * logf(u1));
        x += r * cosf(2 * PI_F * u2);
        y += r * sinf(2 * PI_F * u2);
    } else {
        x += px.half_x * (2 * u1 - 1);
        y += px.half_y * (2 * u2 - 1);
    }
    p.pos = { x, y, px.z };
    p.dir = P.dir;
    p.e = px.energy;
    tag = px.tally;
}

// Occupancy.  The loop is a chain of dependent loads and issue-bound, so
// more resident wavefronts hide more of it: the table variants are built for
// 5 wavefronts per SIMD (96 VGPRs; 5 live values spill to scratch around the
// rare paths), 6.6 % faster than 4 without spills (and 6 per SIMD, 80 VGPRs
// with 17 spilled, is 14 % slower than 4).  Their LDS (stack + path memo of
// 20 wavefronts) fits 160 KB for stack depths up to 112.
// LEARNED: the elastic and inelastic ICDF draws come from the learned
// sampler's networks (evaluated by the whole wavefront) instead of the tables.
// The learned variant needs ~150 registers while the networks run; it is
// built for 4 wavefronts per SIMD all the same (128 VGPRs, 9 of them spilled
// around the networks: 10 % faster than 3 wavefronts per SIMD without
// spills).  With the deposit tally too it would spill 19, so that variant is
// built for 3 wavefronts per SIMD.
// DEPOSIT: the energy-deposition tally is on (a separate instantiation, so
// that the plain kernel carries none of its registers).  PENN: the material
// set's inelastic model is full Penn (its event runs on its own, not merged
// with the elastic lanes' direction frame).  WORK: the work counters
// (events, table lookups, node steps, triangle tests, boundary events,
// secondaries, waits) are kept; without them a wavefront counts only
// primaries, detections and stopped electrons.  The counters are a dozen
// wave-uniform sums and two per-lane ones updated every loop iteration, and
// cost about 4 % of the kernel time (more scalar registers than the loop
// has, so they live in vector-register lanes), so runs that do not report
// them skip them.  The deposit kernels always count.
template<bool LEARNED, bool DEPOSIT, bool PENN, bool WORK>
__global__ __launch_bounds__(LEARNED && DEPOSIT ? 768 : 1024) __attribute__((amdgpu_waves_per_eu(LEARNED ? 1 : 5)))
void k_persistent(pers_params P)
{
    extern __shared__ float lds[];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const uint32_t gwave = blockIdx.x * (blockDim.x >> 6) + wave;
    const wave_stack st{ (lds_float*)lds + (size_t)wave * 9 * P.depth, P.overflow + (size_t)gwave * 9 * P.ovf,
                         P.depth, P.ovf };
    const uint32_t cap = P.depth + P.ovf;
    // per-lane octree path memo after all stacks: [wave][level][lane]
    uint32_t* path_lds = reinterpret_cast<uint32_t*>(lds + (size_t)(blockDim.x >> 6) * 9 * P.depth) +
                         (size_t)wave * lds_path_cache::kWords * 64 + lane;
    lds_path_cache pcache{ path_lds, 1ull, 0ull, 0u };

    rng_xorwow rng{ rng_seed(P.rng_seed, P.rng_first + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x) };
    dparticle p{ { 0, 0, 0 }, { 0, 0, 1 }, 0 };
    int mat = MAT_VACUUM;
    uint32_t last = NO_TRI, tag = 0, steps = 0;
    bool alive = false, pend = false;
    // a step's walk may span several loop iterations, one cell per iteration
    bool walking = false;
    walk_state w = walk_begin(p.pos, 0);
    float dist = 0;
    uint8_t type = 0;
    uint32_t sp = 0;                                // wave-uniform stack size
    uint32_t waited = 0;                            // iterations an inelastic request has waited (wave-uniform)
    const uint32_t home = blockIdx.x % N_QUEUES;    // this block's XCD
    uint32_t qi = 0;                                // queues found empty so far (wave-uniform)
    // Work counters: per-lane walk counts, and wave-uniform counts of lanes
    // (population counts of ballots, held in scalar registers).
    prop_counters walked{ 0, 0 };
    uint32_t n_prims = 0, n_events = 0, n_l1d = 0, n_l2d = 0, n_rd = 0, n_bound = 0, n_sec = 0, n_pend = 0,
             n_se = 0, n_bse = 0, n_killed = 0, n_dep = 0;
    double dep_sum = 0;                               // this lane's deposited energy

    if (!LEARNED) __builtin_amdgcn_s_setprio(2);
    for (;;) {
        // ---- refill idle lanes: stack first, then fresh primaries ----------
        // (a full stack refills at once: popping is what makes room for
        // the secondaries of waiting inelastic events)
        const uint64_t idle = __ballot(!alive);
        const uint32_t n_idle = __popcll(idle);
        if (n_idle && (n_idle >= P.refill_threshold || n_idle == 64 || sp == cap)) {
            const uint32_t take = min(n_idle, sp);
            const uint32_t r = rank_below(idle);
            if (!alive && r < take) {
                st.get(sp - 1 - r, p, mat, tag);
                last = NO_TRI; steps = 0; pend = false; alive = true;
            }
            sp -= take;
            wave_sync();
            const uint32_t need = n_idle - take;
            // fresh primaries: the home queue first, then the others in turn
            while (need && qi < N_QUEUES) {
                const uint32_t q = (home + qi) % N_QUEUES;
                const uint64_t qb = P.qbeg[q], qe = P.qbeg[q + 1];
                uint64_t base = 0;
                if (lane == 0) base = atomicAdd(&P.next[q * QSTRIDE], (unsigned long long)need);
                base = qb + bcast64(base);
                if (base >= qe) { ++qi; continue; }
                if (base + need >= qe) ++qi;
                const uint64_t fresh = __ballot(!alive);
                if (!alive) {
                    const uint64_t i = base + rank_below(fresh);
                    if (i < qe) {
                        make_primary(P, i, p, tag);
                        mat = MAT_VACUUM; last = NO_TRI; steps = 0; pend = false; alive = true;
                    }
                }
                n_prims += (uint32_t)min<uint64_t>(qe - base, (uint64_t)__popcll(fresh));
                break;
            }
        }
        if (__ballot(alive) == 0) break;

        // ---- init step: sample; walk one cell; move when the walk ends ------
        // A lane whose walk needs more cells keeps walking next iteration,
        // while the other lanes go on with their events and next steps, so
        // every iteration walks at most one cell per lane.
        uint8_t ev = EV_NONE;
        bool started = false, capped = false, init = false;
        if (alive && !walking) {
            if (pend) ev = EV_INELASTIC;
            else if (++steps > STEP_CAP) { alive = false; capped = true; }
            else init = true;
        }
        {
            // the sampling lanes' material, read once for the wavefront
            // when they share one
            const int mu = wave_material(init && mat >= 0, mat);
            lane_counters unused;
            bool ok = false;
            if (mu >= 0) {
                const dmat m = load_material(const_mats(P.mats), mu);
                if (init) ok = init_sample_in<false>(P.g, m, mat >= 0, P.min_e, P.max_e, p, rng, unused, dist, type);
            } else {
                if (init) ok = init_sample<false>(P.g, P.mats, P.min_e, P.max_e, p, mat, rng, unused, dist, type);
            }
            if (init) {
                if (ok) { started = walking = true; w = walk_begin(p.pos, dist); }
                else alive = false;
            }
        }
        n_killed += __popcll(__ballot(capped));
        if (WORK) {
            n_events += __popcll(__ballot(init));
            n_l1d += 2 * __popcll(__ballot(started && mat >= 0));
        }
        // The cell step runs at the highest wavefront priority (3), refills
        // and path sampling at 2, the events at 0: the first two are chains
        // of dependent loads, and where several wavefronts of a SIMD are
        // ready, the one issuing its next load first overlaps the others'
        // arithmetic (-0.8 % kernel time, measured; the events raised above
        // the refills were slower, and the learned kernel, whose events hold
        // the networks' matrix products, runs without priorities).
        if (!LEARNED) __builtin_amdgcn_s_setprio(3);
        {
            // all lanes enter the cell step; lanes not walking help test triangles
            uint32_t hit = NO_TRI;
            float d = 0;
            if (walk_cell<WORK>(walking, P.g, w, p.dir, dist, last, mat, &hit, &d, &walked, &pcache)) {
                walking = false;
                ev = init_move(p, dist, d, hit, type, last);
            }
        }
        if (!LEARNED) __builtin_amdgcn_s_setprio(0);
        // ---- inelastic events need room for a secondary --------------------
        // A wavefront whose live lanes all wait on a full stack (no lane is
        // idle: an idle lane would have popped an entry above) goes ahead
        // anyway and drops (counts) the secondaries that do not fit, so that
        // every wavefront keeps making progress; the host reports the drops.
        const bool want = alive && ev == EV_INELASTIC;
        const uint64_t wm = __ballot(want), am = __ballot(alive);
        const bool batch = (uint32_t)__popcll(wm) >= P.inel_batch || waited >= P.inel_wait || wm == am;
        waited = (wm && !batch) ? waited + 1 : 0;
        const uint32_t room = cap - sp;
        const bool stuck = room == 0 && wm == am;
        const bool go = want && batch && (stuck || rank_below(wm) < room);
        if (want && !go) pend = true;
        if (WORK) {
            n_pend += __popcll(__ballot(want && batch && !go));
            n_bound += __popcll(__ballot(alive && ev == EV_INTERSECT));
            const uint32_t n_el = __popcll(__ballot(alive && ev == EV_ELASTIC)), n_in = __popcll(__ballot(go));
            n_l2d += n_el + n_in;
            n_rd += n_in;
        }

        // ---- event ----------------------------------------------------------
        bool made = false, se = false, bse = false;
        dparticle q;
        if (alive) {
            if (ev == EV_INTERSECT) {
                const tri_rec t = load_tri(P.g, last);
                const int res = intersect_execute(P.mats, t, p, mat, rng);
                if (res == OUT_DETECTED) {
                    alive = false;
                    if (p.e < 50) { se = true; if (P.pixel_se) atomicAdd(&P.pixel_se[tag], 1u); }
                    else { bse = true; if (P.pixel_bse) atomicAdd(&P.pixel_bse[tag], 1u); }
                    const int bin = (int)p.e;
                    if (bin >= 0 && bin < NBL_HIST_BINS) atomicAdd(&P.hist[bin], 1ull);
                } else if (res == OUT_TERMINATED) {
                    alive = false;
                }
            }
        }
        {
            // elastic and inelastic lanes share one pass through each part
            const bool el = alive && ev == EV_ELASTIC;
            const float e_before = p.e;
            bool rot = false;
            scatter_state ss;
            if (LEARNED) {
                // the networks of P.learned (wave-uniform) replace their
                // tables; a kind left out is read from its table as usual
                float y = 0.5f, le = 0.0f;
                if (el || go) { y = rng.unit(); le = log_fast(p.e); }
                float ve = 0, vi = 0;
                if (P.learned & (1u << MLP_KIND_ELASTIC)) ve = mlp_eval_needed(P.mlp[MLP_KIND_ELASTIC], el, mat, le, y);
                else if (el) ve = tab2_get(P.mats[mat].el_icdf, le, y);
                if (P.learned & (1u << MLP_KIND_INELASTIC)) vi = mlp_eval_needed(P.mlp[MLP_KIND_INELASTIC], go, mat, le, y);
                else if (go) vi = tab2_get(P.mats[mat].in_icdf, le, y);
                const float v = el ? ve : vi;
                if (el || go) {
                    pend = false;
                    last = NO_TRI;
                    rot = scatter_begin_v(P.mats[mat], p, rng, el, v, ss);
                }
            }
            // Kieft events from the tables: the scattering lanes' material is
            // read once for the wavefront when they share one
            auto kieft = [&](const dmat& m) {
                if (!LEARNED && (el || (go && !PENN))) {
                    pend = false;
                    last = NO_TRI;
                    rot = scatter_begin(m, p, rng, el, ss);
                }
                f3 nd{ 0, 0, 0 };
                if (rot) nd = scatter_frame(p, rng);
                if (rot) made = scatter_end(m, p, rng, ss, nd, &q);
            };
            const int mu = LEARNED || PENN ? -1 : wave_material(el || go, mat);
            if (mu >= 0) {
                const dmat m = load_material(const_mats(P.mats), mu);
                kieft(m);
            } else {
                kieft(P.mats[mat >= 0 ? mat : 0]);
            }
            if (PENN && go) {
                pend = false;
                last = NO_TRI;
                made = penn_execute(P.mats[mat], p, rng, &q);
            }
            // energy deposit of a scatter event (energydeposit_cpu_driver.h:72-94)
            if (DEPOSIT) {
                const float de = (el || go) ? e_before - p.e : 0.0f;
                const bool dep = de != 0.0f;
                const uint64_t dm = __ballot(dep);
                if (WORK) n_dep += __popcll(dm);
                if (dep) {
                    dep_sum += de;
                    const int ix = (int)floorf((p.pos.x - P.dlo.x) * P.dinv.x);
                    const int iy = (int)floorf((p.pos.y - P.dlo.y) * P.dinv.y);
                    const int iz = (int)floorf((p.pos.z - P.dlo.z) * P.dinv.z);
                    if (P.dep && ix >= 0 && iy >= 0 && iz >= 0 && ix < (int)P.dnx && iy < (int)P.dny &&
                        iz < (int)P.dnz)
                        atomicAdd(&P.dep[((size_t)iz * P.dny + iy) * P.dnx + ix], (double)de);
                }
                // records as cpu_energydep.cpp:184-192 writes them: one counter update per wavefront
                if (P.drec_on && dm) {
                    unsigned long long base = 0;
                    if (lane == 0) base = atomicAdd(P.drec_count, (unsigned long long)__popcll(dm));
                    base = bcast64(base);
                    if (dep) {
                        const uint64_t k = base + rank_below(dm);
                        if (k < P.drec_cap) { P.drec[k] = drecord{ p.pos, e_before, de }; P.drec_tags[k] = tag; }
                    }
                }
            }
        }
        n_se += __popcll(__ballot(se));
        n_bse += __popcll(__ballot(bse));
        // ---- detection records: one counter update per wavefront -------------
        if (P.det_on) {
            const uint64_t dm = __ballot(se || bse);
            if (dm) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(P.det_count, (unsigned long long)__popcll(dm));
                base = bcast64(base);
                if (se || bse) {
                    const uint64_t k = base + rank_below(dm);
                    if (k < P.det_cap) { P.det[k] = p; P.det_tags[k] = tag; }
                }
            }
        }
        // ---- push secondaries ------------------------------------------------
        const uint64_t mm = __ballot(made);
        const uint32_t n_made = __popcll(mm);
        if (made) {
            const uint32_t j = sp + rank_below(mm);
            if (j < cap) st.put(j, q, mat, tag);
        }
        if (WORK) n_sec += n_made;
        n_killed += sp + n_made > cap ? sp + n_made - cap : 0u;   // dropped: no room left
        sp = min(sp + n_made, cap);
        wave_sync();
        if (!LEARNED) __builtin_amdgcn_s_setprio(2);
    }

    // ---- counters: one atomic per wavefront per counter ----------------------
    const uint32_t wave_vals[N_COUNTERS] = { n_prims, n_se, n_bse, n_events, n_l1d, n_l2d, n_rd, 0, 0, n_bound,
                                             n_sec, n_pend, n_killed, n_se + n_bse, n_dep };
    for (int k = 0; k < N_COUNTERS; ++k) {
        uint64_t v = wave_vals[k];
        if (k == C_NODES || k == C_TRIS) {
            v = k == C_NODES ? walked.node_steps : walked.tri_tests;
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        }
        if (lane == 0 && v) atomicAdd(&P.counters[k], (unsigned long long)v);
    }
    if (DEPOSIT && P.dep_total) {
        for (int off = 32; off > 0; off >>= 1) dep_sum += __shfl_xor(dep_sum, off);
        if (lane == 0 && dep_sum != 0) atomicAdd(P.dep_total, dep_sum);
    }
}

// Where the primaries of a persistent run come from: a pixel batch cut into
// kernel pixels (kp, offsets), or an explicit list of particles with tags.
struct primary_source {
    std::vector<dpixel> kp;
    std::vector<uint64_t> offsets{ 0 };
    uint32_t n_tally = 0;               // per-pixel tally entries (the caller's pixels)
    f3 dir{ 0, 0, -1 };
    uint64_t seed = 0;
    const nbl_particle* list = nullptr;
    const uint32_t* tags = nullptr;
    uint64_t n_list = 0;
};

// Launch shape and cadence of a persistent run (nbl_persistent_opts; 0 = default).
struct launch_shape {
    int cus = 0;
    uint32_t block = 128, per_cu = 10, depth = 96, thresh = 4, ovf = 16384, ibatch = 8, iwait = 8;
    bool ovf_given = false, allow_drops = false, penn = false;
    bool work = false;                  // keep the work counters (the deposit kernels always do)
    size_t lds = 0;
    size_t n_waves() const { return (size_t)cus * per_cu * (block / 64); }
    uint64_t n_threads() const { return (uint64_t)cus * per_cu * block; }
};

launch_shape resolve_shape(nbl_driver d, const nbl_persistent_opts* opts, bool deposit)
{
    launch_shape L;
    int dev = 0;
    NBL_HIP(hipGetDevice(&dev));
    NBL_HIP(hipDeviceGetAttribute(&L.cus, hipDeviceAttributeMultiprocessorCount, dev));
    // defaults: the table variants 128 x 10 (5 wavefronts per SIMD), stack 96;
    // the learned one 256 x 4 (4 per SIMD), or 256 x 3 with the deposit tally
    // (3 per SIMD), stack 128
    const bool three_waves = d->sampler && deposit;
    if (d->sampler) { L.block = 256; L.per_cu = three_waves ? 3u : 4u; L.depth = 128; }
    if (opts && opts->block_size) L.block = opts->block_size;
    if (opts && opts->blocks_per_cu) L.per_cu = opts->blocks_per_cu;
    if (opts && opts->stack_depth) L.depth = opts->stack_depth;
    if (opts && opts->refill_threshold) L.thresh = opts->refill_threshold;
    L.ovf_given = opts && opts->overflow_depth;
    if (L.ovf_given) L.ovf = opts->overflow_depth;
    if (opts && opts->inelastic_batch) L.ibatch = opts->inelastic_batch;
    if (opts && opts->inelastic_wait) L.iwait = opts->inelastic_wait;
    L.allow_drops = opts && opts->allow_drops;
    L.work = (opts && opts->count_work) || deposit;
    require(L.ibatch <= 64, "nbl_driver_run_persistent: inelastic_batch must be 1..64");
    require(L.iwait <= 1024, "nbl_driver_run_persistent: inelastic_wait must be <= 1024");
    require(L.ovf <= (1u << 20), "nbl_driver_run_persistent: overflow_depth must be <= 2^20");
    require(L.block % 64 == 0 && L.block >= 64 && L.block <= 1024,
        "nbl_driver_run_persistent: block_size must be 64..1024, multiple of 64");
    require(!three_waves || L.block <= 768,
        "nbl_driver_run_persistent: block_size must be <= 768 with a learned sampler and a deposit tally");
    L.penn = materials_view(d->mats, 0).model == INEL_PENN;
    require(!(L.penn && d->sampler), "nbl_driver_run_persistent: the learned sampler covers Kieft inelastic tables only");
    require(L.per_cu >= 1 && L.per_cu <= 32, "nbl_driver_run_persistent: blocks_per_cu must be 1..32");
    require(L.depth >= 64 && L.depth <= 1024, "nbl_driver_run_persistent: stack_depth must be 64..1024");
    require(L.thresh >= 1 && L.thresh <= 64, "nbl_driver_run_persistent: refill_threshold must be 1..64");
    L.lds = (size_t)(L.block / 64) * (9 * L.depth + lds_path_cache::kWords * 64) * sizeof(float);
    require(L.lds <= 160 * 1024, "nbl_driver_run_persistent: stack_depth x waves exceeds the LDS size");
    return L;
}

// The parameters every persistent launch shares (geometry, materials, shape,
// RNG streams); the caller sets the primaries, tallies and scratch pointers.
pers_params base_params(nbl_driver d, const launch_shape& L)
{
    pers_params P{};
    P.g = geometry_view(d->geom);
    P.mats = materials_device(d->mats);
    P.min_e = d->min_energy; P.max_e = d->max_energy;
    P.rng_seed = d->seed;
    P.rng_first = (uint64_t)d->cap + (d->runs++) * L.n_threads();   // streams distinct from the driver's slots
    P.depth = L.depth; P.ovf = L.ovf; P.refill_threshold = L.thresh;
    P.inel_batch = L.ibatch; P.inel_wait = L.iwait;
    if (d->sampler) {
        const dmlp* v = sampler_views(d->sampler);
        P.mlp[0] = v[0]; P.mlp[1] = v[1];
        P.learned = sampler_kinds(d->sampler);
    }
    return P;
}

// Launches the transport on the driver's stream between the driver's two
// events and returns the launch's device time in ms (waits for it).
float launch_transport(nbl_driver d, const launch_shape& L, const pers_params& P, bool deposit)
{
    hipStream_t s = d->stream;
    NBL_HIP(hipEventRecord(d->ev_start, s));
    const dim3 grid((unsigned)(L.cus * L.per_cu)), threads_per_block(L.block);
    if (d->sampler) {
        if (deposit) hipLaunchKernelGGL((k_persistent<true, true, false, true>), grid, threads_per_block, L.lds, s, P);
        else if (L.work) hipLaunchKernelGGL((k_persistent<true, false, false, true>), grid, threads_per_block, L.lds, s, P);
        else hipLaunchKernelGGL((k_persistent<true, false, false, false>), grid, threads_per_block, L.lds, s, P);
    } else if (L.penn) {
        if (deposit) hipLaunchKernelGGL((k_persistent<false, true, true, true>), grid, threads_per_block, L.lds, s, P);
        else if (L.work) hipLaunchKernelGGL((k_persistent<false, false, true, true>), grid, threads_per_block, L.lds, s, P);
        else hipLaunchKernelGGL((k_persistent<false, false, true, false>), grid, threads_per_block, L.lds, s, P);
    } else {
        if (deposit) hipLaunchKernelGGL((k_persistent<false, true, false, true>), grid, threads_per_block, L.lds, s, P);
        else if (L.work) hipLaunchKernelGGL((k_persistent<false, false, false, true>), grid, threads_per_block, L.lds, s, P);
        else hipLaunchKernelGGL((k_persistent<false, false, false, false>), grid, threads_per_block, L.lds, s, P);
    }
    NBL_HIP(hipGetLastError());
    NBL_HIP(hipEventRecord(d->ev_stop, s));
    NBL_HIP(hipEventSynchronize(d->ev_stop));
    float ms = 0;
    NBL_HIP(hipEventElapsedTime(&ms, d->ev_start, d->ev_stop));
    return ms;
}

// Counters of a finished launch into *out; a run that stopped electrons is
// an error unless the caller allowed drops.
void finish_tallies(const uint64_t* cnt, const unsigned long long* hist, float ms, const launch_shape& L,
    nbl_tallies* out)
{
    if (out->hist) for (int b = 0; b < NBL_HIST_BINS; ++b) out->hist[b] = (uint32_t)hist[b];
    out->primaries = cnt[C_PRIM]; out->detected_se = cnt[C_SE]; out->detected_bse = cnt[C_BSE];
    out->events = cnt[C_EVENTS]; out->lookups_1d = cnt[C_L1D]; out->lookups_2d = cnt[C_L2D];
    out->rounddowns = cnt[C_RD]; out->node_steps = cnt[C_NODES]; out->tri_tests = cnt[C_TRIS];
    out->boundary_events = cnt[C_BOUND]; out->secondaries = cnt[C_SEC]; out->pending_retries = cnt[C_PEND];
    out->stopped = cnt[C_KILLED];
    out->kernel_ms = ms;
    if (out->stopped && !L.allow_drops)
        throw std::runtime_error("nbl_driver_run_persistent: " + std::to_string(out->stopped) +
            " electrons stopped (secondaries dropped for lack of stack room, or the 2^20-step cap); the tallies "
            "are biased. Raise overflow_depth, or set allow_drops to accept them");
}

void run_persistent(nbl_driver d, const primary_source& src, const nbl_persistent_opts* opts, nbl_tallies* out)
{
    const uint64_t total = src.list ? src.n_list : src.offsets.back();
    const uint32_t nkp = (uint32_t)src.kp.size();
    const uint32_t np = src.n_tally;
    require(total > 0, "nbl_driver_run_persistent: no primaries in the batch");
    const launch_shape L = resolve_shape(d, opts, out->deposit != nullptr || out->deposits != nullptr);

    dbuf<dpixel> dpix(std::max<uint32_t>(nkp, 1));
    dbuf<uint64_t> doff(nkp + 1);
    dbuf<unsigned long long> dnext(N_QUEUES * QSTRIDE), dhist(NBL_HIST_BINS), dcnt(N_COUNTERS);
    dbuf<uint32_t> dse(std::max<uint32_t>(np, 1)), dbse(std::max<uint32_t>(np, 1));
    hipStream_t s = d->stream;
    if (nkp) {
        dpix.upload(src.kp.data(), nkp, s);
        doff.upload(src.offsets.data(), nkp + 1, s);
    }
    // explicit primaries and detection records
    dbuf<dparticle> dlist, ddet;
    dbuf<uint32_t> dltags, ddtags;
    dbuf<unsigned long long> ddcnt(1);
    NBL_HIP(hipMemsetAsync(ddcnt.p, 0, sizeof(unsigned long long), s));
    if (src.list) {
        static_assert(sizeof(dparticle) == sizeof(nbl_particle), "particle layout");
        dlist.alloc(total); dltags.alloc(total);
        dlist.upload(reinterpret_cast<const dparticle*>(src.list), total, s);
        dltags.upload(src.tags, total, s);
    }
    nbl_detection_buffer* db = out->detected;
    if (db) {
        require(db->capacity == 0 || (db->particles && db->tags),
            "nbl_driver_run_persistent: detection buffer needs particle and tag arrays");
        if (db->capacity) { ddet.alloc(db->capacity); ddtags.alloc(db->capacity); }
    }
    NBL_HIP(hipMemsetAsync(dnext.p, 0, N_QUEUES * QSTRIDE * sizeof(unsigned long long), s));
    NBL_HIP(hipMemsetAsync(dhist.p, 0, NBL_HIST_BINS * sizeof(unsigned long long), s));
    NBL_HIP(hipMemsetAsync(dcnt.p, 0, N_COUNTERS * sizeof(unsigned long long), s));
    NBL_HIP(hipMemsetAsync(dse.p, 0, std::max<uint32_t>(np, 1) * sizeof(uint32_t), s));
    NBL_HIP(hipMemsetAsync(dbse.p, 0, std::max<uint32_t>(np, 1) * sizeof(uint32_t), s));

    pers_params P = base_params(d, L);
    P.pixels = dpix.p; P.offsets = doff.p; P.n_pixels = nkp; P.total = total;
    P.plist = src.list ? dlist.p : nullptr; P.ptags = src.list ? dltags.p : nullptr;
    P.det_on = db ? 1u : 0u;             // capacity 0: count only
    P.det = ddet.p; P.det_tags = ddtags.p;
    P.det_count = ddcnt.p; P.det_cap = db ? db->capacity : 0;
    for (uint32_t q = 0; q <= N_QUEUES; ++q) P.qbeg[q] = total * q / N_QUEUES;   // contiguous eighths
    P.dir = src.dir;
    P.prim_seed = src.seed;
    P.next = dnext.p; P.hist = dhist.p; P.counters = dcnt.p;
    P.pixel_se = np ? dse.p : nullptr; P.pixel_bse = np ? dbse.p : nullptr;
    const size_t ovf_floats = L.n_waves() * 9 * L.ovf;
    if (d->overflow.n < ovf_floats) d->overflow.alloc(ovf_floats);
    P.overflow = d->overflow.p;
    // energy-deposition tally
    nbl_deposit_grid* dg = out->deposit;
    dbuf<double> ddep, dtot;
    P.dep = nullptr; P.dep_total = nullptr;
    if (dg) {
        require(dg->energy != nullptr && dg->nx > 0 && dg->ny > 0 && dg->nz > 0,
            "nbl_driver_run_persistent: deposit grid needs nx, ny, nz > 0 and an energy array");
        require((uint64_t)dg->nx * dg->ny * dg->nz <= (1ull << 30), "nbl_driver_run_persistent: deposit grid too large");
        require(dg->hi.x > dg->lo.x && dg->hi.y > dg->lo.y && dg->hi.z > dg->lo.z,
            "nbl_driver_run_persistent: deposit grid box is empty");
        const size_t nv = (size_t)dg->nx * dg->ny * dg->nz;
        ddep.alloc(nv); dtot.alloc(1);
        NBL_HIP(hipMemsetAsync(ddep.p, 0, nv * sizeof(double), s));
        NBL_HIP(hipMemsetAsync(dtot.p, 0, sizeof(double), s));
        P.dep = ddep.p; P.dep_total = dtot.p;
        P.dnx = dg->nx; P.dny = dg->ny; P.dnz = dg->nz;
        P.dlo = { dg->lo.x, dg->lo.y, dg->lo.z };
        P.dinv = { (float)dg->nx / (dg->hi.x - dg->lo.x), (float)dg->ny / (dg->hi.y - dg->lo.y),
                   (float)dg->nz / (dg->hi.z - dg->lo.z) };
    }

    // energy-deposition records
    nbl_deposit_buffer* drb = out->deposits;
    dbuf<drecord> ddrec;
    dbuf<uint32_t> ddrtags;
    dbuf<unsigned long long> ddrcnt(1);
    NBL_HIP(hipMemsetAsync(ddrcnt.p, 0, sizeof(unsigned long long), s));
    P.drec_on = 0; P.drec = nullptr; P.drec_tags = nullptr; P.drec_count = ddrcnt.p; P.drec_cap = 0;
    if (drb) {
        static_assert(sizeof(drecord) == sizeof(nbl_deposit_record), "deposit record layout");
        require(drb->capacity == 0 || (drb->records && drb->tags),
            "nbl_driver_run_persistent: deposit record buffer needs record and tag arrays");
        if (drb->capacity) { ddrec.alloc(drb->capacity); ddrtags.alloc(drb->capacity); }
        P.drec_on = 1; P.drec = ddrec.p; P.drec_tags = ddrtags.p; P.drec_cap = drb->capacity;
    }

    const float ms = launch_transport(d, L, P, dg != nullptr || drb != nullptr);

    uint64_t cnt[N_COUNTERS];
    dcnt.download(reinterpret_cast<unsigned long long*>(cnt), N_COUNTERS, s);
    std::vector<unsigned long long> hist(NBL_HIST_BINS);
    dhist.download(hist.data(), NBL_HIST_BINS, s);
    if (out->pixel_se && np) dse.download(out->pixel_se, np, s);
    if (out->pixel_bse && np) dbse.download(out->pixel_bse, np, s);
    unsigned long long n_det = 0;
    ddcnt.download(&n_det, 1, s);
    NBL_HIP(hipStreamSynchronize(s));
    if (db) {
        db->count = n_det;
        const uint64_t k = std::min<uint64_t>(n_det, db->capacity);
        if (k) {
            ddet.download(reinterpret_cast<dparticle*>(db->particles), k, s);
            ddtags.download(db->tags, k, s);
        }
    }
    if (dg) {
        ddep.download(dg->energy, (size_t)dg->nx * dg->ny * dg->nz, s);
        dtot.download(&dg->total, 1, s);
    }
    if (drb) {
        unsigned long long n_rec = 0;
        ddrcnt.download(&n_rec, 1, s);
        NBL_HIP(hipStreamSynchronize(s));
        drb->count = n_rec;
        const uint64_t k = std::min<uint64_t>(n_rec, drb->capacity);
        if (k) {
            ddrec.download(reinterpret_cast<drecord*>(drb->records), k, s);
            ddrtags.download(drb->tags, k, s);
        }
    }
    NBL_HIP(hipStreamSynchronize(s));
    if (dg) dg->count = cnt[C_DEPN];
    finish_tallies(cnt, hist.data(), ms, L, out);
}

// ---- the allocation-free variant's scratch -----------------------------------
// One fixed array per GPU, for callers that cannot allocate during a run:
// the control words (primary queues, work counters, detection count, the
// invalid-primary count, the energy histogram) and the per-wavefront HBM
// secondary stacks.  Host code never passes `scratch` itself to a kernel:
// its device address comes from hipGetSymbolAddress, once per device.
constexpr size_t NA_NEXT = 0;                                  // N_QUEUES x QSTRIDE words
constexpr size_t NA_CNT = NA_NEXT + N_QUEUES * QSTRIDE;        // N_COUNTERS words
constexpr size_t NA_DET = NA_CNT + 16;                         // detections
constexpr size_t NA_BAD = NA_DET + 1;                          // invalid primaries
constexpr size_t NA_HIST = 256;                                // NBL_HIST_BINS words
constexpr size_t NA_CTRL_WORDS = NA_HIST + NBL_HIST_BINS;      // cleared before every run
constexpr size_t NA_STACK_WORDS = (size_t)NBL_NOALLOC_STACK_BYTES / 8;
static_assert(NA_BAD < NA_HIST, "control words overlap the histogram");
static_assert(N_COUNTERS <= 16, "counter block");

static __device__ __align__(256) uint64_t scratch[NA_CTRL_WORDS + NA_STACK_WORDS];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
uint64_t* na_ptr[NA_MAX_DEVICES] = {};

// This device's copy of `scratch` (looked up on the first call per device).
uint64_t* scratch_on_current_device()
{
    int dev = 0;
    NBL_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < NA_MAX_DEVICES, "nbl_driver_run_persistent_list_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void* p = nullptr;
        NBL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "nbl_driver_run_persistent_list_noalloc: no device address for the scratch array");
        na_ptr[dev] = static_cast<uint64_t*>(p);
    }
    return na_ptr[dev];
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint32_t n, uint64_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Counts the primaries the list entry point would refuse: a non-finite
// position, a zero or non-finite direction, a non-finite or non-positive energy.
__global__ void k_count_invalid(const dparticle* __restrict__ q, uint64_t n, unsigned long long* __restrict__ bad)
{
    uint32_t c = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const dparticle p = q[i];
        const float dd = p.dir.x * p.dir.x + p.dir.y * p.dir.y + p.dir.z * p.dir.z;
        const bool ok = isfinite(p.pos.x) && isfinite(p.pos.y) && isfinite(p.pos.z) && isfinite(dd) && dd > 0 &&
                        isfinite(p.e) && p.e > 0;
        c += ok ? 0u : 1u;
    }
    if (c) atomicAdd(bad, (unsigned long long)c);
}

void run_persistent_noalloc(nbl_driver d, const nbl_particle* primaries, const uint32_t* tags, uint64_t n,
    nbl_particle* det, uint32_t* det_tags, uint64_t det_cap, const nbl_persistent_opts* opts, nbl_tallies* out)
{
    require(n > 0 && primaries && tags, "nbl_driver_run_persistent_list_noalloc: no primaries");
    require(!out->pixel_se && !out->pixel_bse,
        "nbl_driver_run_persistent_list_noalloc: per-pixel tallies need a pixel batch (use the detection records)");
    require(!out->deposit && !out->detected && !out->deposits,
        "nbl_driver_run_persistent_list_noalloc: out->deposit, out->detected and out->deposits must be NULL (the "
        "deposit tallies need an allocation; detection records go to the device arrays passed as arguments)");
    require(det_cap == 0 || (det && det_tags),
        "nbl_driver_run_persistent_list_noalloc: detected_capacity > 0 needs both record arrays");
    const launch_shape L0 = resolve_shape(d, opts, false);
    launch_shape L = L0;
    const size_t per_entry = 9 * sizeof(float) * L.n_waves();
    const size_t max_ovf = (size_t)NBL_NOALLOC_STACK_BYTES / per_entry;
    if (!L.ovf_given) L.ovf = (uint32_t)std::min<size_t>(L.ovf, max_ovf);
    if (L.ovf > max_ovf)
        throw invalid_arg("nbl_driver_run_persistent_list_noalloc: overflow_depth " + std::to_string(L.ovf) +
            " x " + std::to_string(L.n_waves()) + " wavefronts exceeds the scratch's " +
            std::to_string(NBL_NOALLOC_STACK_BYTES >> 20) + " MiB of stack: at most " + std::to_string(max_ovf) +
            " entries per wavefront at this launch shape");

    uint64_t* sp = scratch_on_current_device();
    hipStream_t s = d->stream;
    unsigned long long* ctrl = reinterpret_cast<unsigned long long*>(sp);
    // clear the control words with a kernel (no memset node if the caller captures this stream)
    hipLaunchKernelGGL(k_fill_u64, dim3(4), dim3(256), 0, s, sp, (uint32_t)NA_CTRL_WORDS, (uint64_t)0);
    hipLaunchKernelGGL(k_count_invalid, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s,
        reinterpret_cast<const dparticle*>(primaries), n, ctrl + NA_BAD);
    NBL_HIP(hipGetLastError());
    unsigned long long bad = 0;
    NBL_HIP(hipMemcpyAsync(&bad, ctrl + NA_BAD, sizeof bad, hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    require(bad == 0, "nbl_driver_run_persistent_list_noalloc: a primary has a non-finite position, a zero "
                      "direction or no energy");

    pers_params P = base_params(d, L);
    P.pixels = nullptr; P.offsets = nullptr; P.n_pixels = 0; P.total = n;
    P.plist = reinterpret_cast<const dparticle*>(primaries); P.ptags = tags;
    P.det_on = det_cap ? 1u : 0u;
    P.det = reinterpret_cast<dparticle*>(det); P.det_tags = det_tags;
    P.det_count = ctrl + NA_DET; P.det_cap = det_cap;
    for (uint32_t q = 0; q <= N_QUEUES; ++q) P.qbeg[q] = n * q / N_QUEUES;
    P.next = ctrl + NA_NEXT; P.hist = ctrl + NA_HIST; P.counters = ctrl + NA_CNT;
    P.pixel_se = nullptr; P.pixel_bse = nullptr;
    P.overflow = reinterpret_cast<float*>(sp + NA_CTRL_WORDS);
    P.dep = nullptr; P.dep_total = nullptr;
    P.drec_on = 0; P.drec = nullptr; P.drec_tags = nullptr; P.drec_count = nullptr; P.drec_cap = 0;

    const float ms = launch_transport(d, L, P, false);

    uint64_t cnt[N_COUNTERS];
    unsigned long long hist[NBL_HIST_BINS];
    NBL_HIP(hipMemcpyAsync(cnt, ctrl + NA_CNT, sizeof cnt, hipMemcpyDeviceToHost, s));
    if (out->hist) NBL_HIP(hipMemcpyAsync(hist, ctrl + NA_HIST, sizeof hist, hipMemcpyDeviceToHost, s));
    NBL_HIP(hipStreamSynchronize(s));
    finish_tallies(cnt, hist, ms, L, out);
}

}  // namespace

extern "C" {

int nbl_driver_run_persistent(nbl_driver d, const nbl_pixel_batch* batch, const nbl_persistent_opts* opts,
    nbl_tallies* out)
{
    return guarded([&] {
        require(d != nullptr, "nbl_driver_run_persistent: NULL driver");
        require(batch && out, "nbl_driver_run_persistent: NULL batch or tallies");
        require(batch->n_pixels > 0 && batch->pixels, "nbl_driver_run_persistent: empty pixel batch");
        const uint32_t np = batch->n_pixels;
        // Kernel pixel list: caller's pixels in order; a large uniform spot is
        // cut into kTiles x kTiles tiles in raster order (stratified: each tile
        // gets its share of the primaries), so that contiguous primary ranges
        // are spatially compact.
        constexpr uint32_t kTiles = 16;
        std::vector<dpixel> kp;
        std::vector<uint64_t> offsets(1, 0);
        for (uint32_t i = 0; i < np; ++i) {
            const nbl_pixel& px = batch->pixels[i];
            require(px.profile <= 1, "nbl_driver_run_persistent: pixel profile must be 0 or 1");
            require(std::isfinite(px.x) && std::isfinite(px.y) && std::isfinite(px.z) && px.energy > 0 &&
                    std::isfinite(px.sigma) && std::isfinite(px.half_x) && std::isfinite(px.half_y),
                "nbl_driver_run_persistent: pixel position/energy/size invalid");
            if (px.n_primaries == 0) continue;
            if (px.profile == 1 && px.n_primaries >= 64u * kTiles * kTiles) {
                const uint32_t per = px.n_primaries / (kTiles * kTiles), extra = px.n_primaries % (kTiles * kTiles);
                for (uint32_t ty = 0; ty < kTiles; ++ty)
                    for (uint32_t tx = 0; tx < kTiles; ++tx) {
                        const uint32_t k = ty * kTiles + tx;
                        dpixel t{ px.x + px.half_x * ((float)(2 * tx + 1) / kTiles - 1.0f),
                                  px.y + px.half_y * ((float)(2 * ty + 1) / kTiles - 1.0f), px.z, px.energy,
                                  0.0f, px.half_x / kTiles, px.half_y / kTiles, 1u, i };
                        kp.push_back(t);
                        offsets.push_back(offsets.back() + per + (k < extra ? 1 : 0));
                    }
            } else {
                kp.push_back(dpixel{ px.x, px.y, px.z, px.energy, px.sigma, px.half_x, px.half_y, px.profile, i });
                offsets.push_back(offsets.back() + px.n_primaries);
            }
        }
        const f3 dir{ batch->dir.x, batch->dir.y, batch->dir.z };
        const float dn = std::sqrt(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
        require(dn > 0 && std::isfinite(dn), "nbl_driver_run_persistent: beam direction is zero");
        primary_source src;
        src.kp = std::move(kp);
        src.offsets = std::move(offsets);
        src.n_tally = np;
        src.dir = { dir.x / dn, dir.y / dn, dir.z / dn };
        src.seed = batch->seed;
        run_persistent(d, src, opts, out);
    });
}

int nbl_driver_run_persistent_list(nbl_driver d, const nbl_particle* primaries, const uint32_t* tags, uint64_t n,
    const nbl_persistent_opts* opts, nbl_tallies* out)
{
    return guarded([&] {
        require(d != nullptr, "nbl_driver_run_persistent_list: NULL driver");
        require(out != nullptr, "nbl_driver_run_persistent_list: NULL tallies");
        require(n > 0 && primaries && tags, "nbl_driver_run_persistent_list: no primaries");
        for (uint64_t i = 0; i < n; ++i) {
            const nbl_particle& q = primaries[i];
            const float dd = q.dir.x * q.dir.x + q.dir.y * q.dir.y + q.dir.z * q.dir.z;
            require(std::isfinite(q.pos.x) && std::isfinite(q.pos.y) && std::isfinite(q.pos.z) &&
                    std::isfinite(dd) && dd > 0 && std::isfinite(q.kin_energy) && q.kin_energy > 0,
                "nbl_driver_run_persistent_list: a primary has a non-finite position, a zero direction or no energy");
        }
        require(!out->pixel_se && !out->pixel_bse,
            "nbl_driver_run_persistent_list: per-pixel tallies need a pixel batch (use the detection records)");
        primary_source src;
        src.list = primaries; src.tags = tags; src.n_list = n;
        run_persistent(d, src, opts, out);
    });
}

int nbl_driver_run_persistent_list_noalloc(nbl_driver d, const nbl_particle* dev_primaries,
    const uint32_t* dev_tags, uint64_t n, nbl_particle* dev_detected, uint32_t* dev_detected_tags,
    uint64_t detected_capacity, const nbl_persistent_opts* opts, nbl_tallies* out)
{
    return guarded([&] {
        require(d != nullptr, "nbl_driver_run_persistent_list_noalloc: NULL driver");
        require(out != nullptr, "nbl_driver_run_persistent_list_noalloc: NULL tallies");
        run_persistent_noalloc(d, dev_primaries, dev_tags, n, dev_detected, dev_detected_tags, detected_capacity,
            opts, out);
    });
}

int nbl_driver_set_sampler(nbl_driver d, nbl_sampler smp)
{
    return guarded([&] {
        require(d != nullptr, "nbl_driver_set_sampler: NULL driver");
        require(smp == nullptr || sampler_materials(smp) == d->mats,
            "nbl_driver_set_sampler: sampler was trained for another material set");
        d->sampler = smp;
    });
}

}  // extern "C"
