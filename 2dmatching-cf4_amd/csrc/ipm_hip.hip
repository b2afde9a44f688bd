// gfx950 particle hot path of iPM2D: kernels and the C ABI of include/ipm_hip.h.
//
// Layout in HBM: every species owns two buffer sets (current / sort target)
// of 8 FP64 arrays (Z R Vz Vr Vt Az Ar WQ) of its capacity, 256-byte aligned.
// Deleted particles stay in place with Z = kDead until the next cell sort
// compacts them; products of MCC are appended past the slots in use.
//
// One kernel template serves every per-particle entry point.  The fused step
// reads each particle once (64 B) and writes it once (56 B): push, absorb,
// deposit and MCC all happen in registers between the load and the store.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#include <vector>
#include <algorithm>
#include <cmath>
#include <hipcub/block/block_radix_sort.hpp>
#include <hipcub/block/block_scan.hpp>
#include <mutex>
#include "ipm_device.hpp"

using namespace ipm;

namespace {

constexpr int kBlock = 256;          // threads per block (4 wave64) of every kernel but the tuned particle pass
constexpr int kTile = 2048;          // slots per tile: one block of the particle pass
constexpr int kWinZ = 64;            // LDS deposit window (nodes in z)
constexpr int kWinR = 24;            // default LDS deposit window height (nodes in r)
// A tile's window starts this many nodes before its first particle's cell.
constexpr int kAnchorZ = 4;
constexpr int anchor_rows(int winr) { return winr / 2 - 1; }

enum : int { M_PUSH = 1, M_ABSORB = 2, M_DEPOSIT = 4, M_MCC = 8, M_REZONE = 16 };

// Random-stream tag of the merge draw: apart from every (species, gas) tag
// of the collision pass, which stay below 2^16.
constexpr uint32_t kMergeTag = 0x4D520000u;

// A collision-queue record: Z R Vz Vr Vt WQ, the particle's index in its
// tile and a pad, 64 contiguous bytes moved as four 16-byte accesses.  It
// holds everything the selection, the target sampling and the reaction
// choice read, so the collision pass makes no second, dependent read of the
// particle; Az and Ar (read back only for real reactions) stay out.
constexpr int kRecD = 8;
constexpr int kRecParts = kRecD / 2;
typedef double rec2 __attribute__((ext_vector_type(2)));
// Queue record: Z R Vz Vr Vt WQ, index in the tile, and the uniform of the
// reaction choice (draw 1 of the particle's stream; with one gas only).
__device__ __forceinline__ unsigned load_rec(const double *r, Part &p, double &u)
{
    const auto *v = (const __attribute__((address_space(1))) rec2 *)r;
    const rec2 a = v[0], b = v[1], c = v[2], d = v[3];
    p.Z = a.x; p.R = a.y; p.Vz = b.x; p.Vr = b.y; p.Vt = c.x; p.WQ = c.y; p.Az = 0.0; p.Ar = 0.0;
    u = d.y;
    return (unsigned)__double_as_longlong(d.x);
}

__device__ __forceinline__ bool is_dead(double z) { return z < -1.0e299; }

// ------------------------------------------------------------ products
// Products of one collision block, staged in LDS and appended by
// collide_append_kernel without any contended atomic.
constexpr int kProdStage = 256;
constexpr int kCollideBlocksPerCuMax = 6;   // the product stages are sized for this many
struct ProdSink {
    lds_ptr<double> rec;        // kProdStage x 8 doubles
    lds_ptr<unsigned> meta;     // species | rank << 8
    lds_ptr<unsigned> count;    // staged so far
    lds_ptr<unsigned> spc;      // staged per species (IPM_NS_MAX)
};

// twin != 0: the product would be a bit-identical copy of another particle
// of its species made by the same collision (dissociation into two equal
// fragments, 113 giving back the parent's own species); the low bit of its
// Vt pattern is flipped so that the two draw from different state-keyed
// streams afterwards, as split copies do.
__device__ __forceinline__ void emit_product(const DevState &st, DevCounters *ctr, Part q, int d,
                                             const ProdSink *ps = nullptr, bool twin = false)
{
    if (d < 0 || d >= st.ns) {
        global_or(&ctr->overflow, 2);
        return;
    }
    q.Vz = q.Vz * st.inv_vfactor;
    q.Vr = q.Vr * st.inv_vfactor;
    q.Vt = q.Vt * st.inv_vfactor;
    if (twin) q.Vt = __longlong_as_double(__double_as_longlong(q.Vt) ^ 1ll);
    if (ps) {
        const unsigned k = __hip_atomic_fetch_add(ps->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k < (unsigned)kProdStage) {
            const unsigned r = __hip_atomic_fetch_add(ps->spc + d, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            lds_ptr<double> o = ps->rec + (size_t)k * 8;
            o[0] = q.Z; o[1] = q.R; o[2] = q.Vz; o[3] = q.Vr; o[4] = q.Vt; o[5] = q.Az; o[6] = q.Ar; o[7] = q.WQ;
            ps->meta[k] = (unsigned)d | (r << 8);
            return;
        }
    }
    // no sink, or the block's stage is full: reserve a slot directly
    unsigned long long k = global_fetch_add(&ctr->n_new[d], 1ull);
    unsigned long long pos = gp(ctr->n)[d] + k;
    if (pos >= (unsigned long long)st.cap[d]) {
        global_or(&ctr->overflow, 1);
        return;
    }
    double *const *F = st.f[d];
    gp(F[0])[pos] = q.Z; gp(F[1])[pos] = q.R; gp(F[2])[pos] = q.Vz; gp(F[3])[pos] = q.Vr;
    gp(F[4])[pos] = q.Vt; gp(F[5])[pos] = q.Az; gp(F[6])[pos] = q.Ar; gp(F[7])[pos] = q.WQ;
}

// PostVelocityIonReactive, R/src/mc/MCCIon.F90:54-74 (Vy from Vz as written there)
__device__ __forceinline__ void ion_reactive(Mc &m, double th)
{
    double VF = sqrt(1.0 - th / m.energy);
    m.energy = m.energy - th;
    mc_vel_update(m, VF);
    double A = m.mass_ratio, B = 1.0 - A;
    m.poi.Vr = B * m.poi.Vr + A * m.pot.Vr + B * m.gx;
    m.poi.Vt = B * m.poi.Vz + A * m.pot.Vt + B * m.gy;
    m.poi.Vz = B * m.poi.Vz + A * m.pot.Vz + B * m.gz;
}

// One selected particle of species s against gas table t: the body of the
// MCC loop, R/src/mc/MCCPublic.F90:61-97, with SelectCollisionElectron
// (R/src/mc/MCCElectron.F90:13-147) and SelectCollisionIon (R/src/mc/MCCIon.F90:6-51),
// in two halves.  mcc_begin: VelRes(VFactor), Updater and the reaction
// choice with the uniform u the caller drew before it (draw 1 of the
// particle's stream for this gas; Updater's target sampling starts at draw
// 2); it returns the reaction index (0 = null event).
__device__ __forceinline__ int mcc_begin(const DevState &st, int s, const DevTable &t, const Part &p, Mc &m, Rng &rg,
                                         double u)
{
    m.poi = p;
    m.poi.Vz = m.poi.Vz * st.vfactor;
    m.poi.Vr = m.poi.Vr * st.vfactor;
    m.poi.Vt = m.poi.Vt * st.vfactor;
    mc_update(m, st.mass[s], t, rg);
    int idx = 0;
    if (t.model == 1 || t.model == 2) idx = select_probability(m, t, u);
    else if (t.model == 3) idx = select_probability_reactive(m, t, u, rg);
    return idx;
}

// mcc_finish: the collision of reaction idx (none for idx 0), the event
// count and VelRes(1/VFactor) back into p.
__device__ __forceinline__ void mcc_finish(const DevState &st, DevCounters *ctr, int s, int g, const DevTable &t,
                                           Mc &m, Rng &rg, int idx, Part &p, const ProdSink *ps = nullptr)
{
    if (idx > 0) {
        const int ty = gp(t.rtype)[idx - 1];
        const int res = gp(t.resultant)[idx - 1];
        const double th = gp(t.threshold)[idx - 1];
        if (t.model == 1) {
            switch (ty) {
            case 11:
                post_collision(m, true, rg);
                break;
            case 12: {
                double VF = sqrt(1.0 - th / m.energy);
                m.energy = m.energy - th;
                mc_vel_update(m, VF);
                post_collision(m, true, rg);
            } break;
            case 13: {
                Part solid = m.poi;
                double E0 = m.energy;
                double EnergyTemp = m.energy - th;
                double R = rg.next();
                double EnergyOld = 10.0 * tan(R * atan(EnergyTemp) / 20.0);
                if (fabs(EnergyOld) < kMINREAL) EnergyOld = kMINREAL;
                double VF = sqrt(EnergyOld / m.energy);
                mc_vel_update(m, VF);
                post_collision(m, true, rg);
                double EnergyNew = EnergyTemp - EnergyOld;
                double VF2 = sqrt(EnergyNew / E0);
                solid.Vz = solid.Vz * VF2;
                solid.Vr = solid.Vr * VF2;
                solid.Vt = solid.Vt * VF2;
                emit_product(st, ctr, solid, s, ps);
                emit_product(st, ctr, m.pot, res, ps);
            } break;
            case 14:
                emit_product(st, ctr, m.pot, res, ps);
                break;
            case 15: {
                double VF = sqrt(1.0 - th / m.energy);
                m.energy = m.energy - th;
                mc_vel_update(m, VF);
                post_collision(m, true, rg);
                emit_product(st, ctr, m.pot, res % 10, ps);
                emit_product(st, ctr, m.pot, res / 10, ps, res % 10 == res / 10);
            } break;
            default:
                break;
            }
        } else {
            switch (ty) {
            case 101: post_collision(m, false, rg); break;
            case 102:
            case 103: m.poi = m.pot; break;
            case 111: ion_reactive(m, th); break;
            case 112:
                ion_reactive(m, th);
                emit_product(st, ctr, m.pot, res, ps);
                break;
            case 113:
                ion_reactive(m, th);
                emit_product(st, ctr, m.poi, res, ps, res == s);
                break;
            default: break;   // 100 and -1 change nothing
            }
        }
    }
    if (st.count_events)
        global_fetch_add(&st.mcc_counts[((size_t)s * st.ng + g) * 128 + (idx & 127)], 1ull);
    m.poi.Vz = m.poi.Vz * st.inv_vfactor;
    m.poi.Vr = m.poi.Vr * st.inv_vfactor;
    m.poi.Vt = m.poi.Vt * st.inv_vfactor;
    p = m.poi;
}

__device__ __forceinline__ void mcc_collide(const DevState &st, DevCounters *ctr, int s, int g,
                                            const DevTable &t, Part &p, Rng &rg, const ProdSink *ps = nullptr)
{
    Mc m;
    const double u = rg.next();
    const int idx = mcc_begin(st, s, t, p, m, rg, u);
    mcc_finish(st, ctr, s, g, t, m, rg, idx, p, ps);
}

__device__ __forceinline__ void lds_add(double *a, double v)
{
    __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}


// DPP row shift right by D within rows of 16 lanes; lanes without a source
// in their row receive 0.
template <int D>
__device__ __forceinline__ double dpp_shr(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x110 + D, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + D, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int D>
__device__ __forceinline__ int dpp_shr_i(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x110 + D, 0xf, 0xf, false);
}
// DPP row broadcasts: CTRL 0x142 (row_bcast:15) gives rows 1 and 3 the last
// lane of the row before them, CTRL 0x143 (row_bcast:31) gives rows 2 and 3
// lane 31; ROWS masks the receiving rows, every other lane receives 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_bcast(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int D>
__device__ __forceinline__ void seg_step(double (&v)[4], int &f)
{
    const double a0 = dpp_shr<D>(v[0]), a1 = dpp_shr<D>(v[1]), a2 = dpp_shr<D>(v[2]), a3 = dpp_shr<D>(v[3]);
    const int fu = dpp_shr_i<D>(f);
    if (!f) {
        v[0] = v[0] + a0; v[1] = v[1] + a1; v[2] = v[2] + a2; v[3] = v[3] + a3;
    }
    f |= fu;
}

template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_bcast(double (&v)[4], int &f)
{
    const double a0 = dpp_bcast<CTRL, ROWS>(v[0]), a1 = dpp_bcast<CTRL, ROWS>(v[1]);
    const double a2 = dpp_bcast<CTRL, ROWS>(v[2]), a3 = dpp_bcast<CTRL, ROWS>(v[3]);
    const int fu = __builtin_amdgcn_update_dpp(0, f, CTRL, ROWS, 0xf, false);
    if (!f) {
        v[0] = v[0] + a0; v[1] = v[1] + a1; v[2] = v[2] + a2; v[3] = v[3] + a3;
    }
    f |= fu;
}

// Segmented inclusive scan over the wave (all 64 lanes active): after the
// call, the last lane of each run of equal keys holds the run's sums.  Any
// key sequence is handled; cell-sorted particles give a few long runs, so
// only a few lanes per wave touch memory.  Returns true on run-end lanes.
__device__ __forceinline__ bool wave_segmented_sum(int key, double (&v)[4])
{
    const int lane = threadIdx.x & 63;
    const int kprev = __shfl_up(key, 1);
    const int knext = __shfl_down(key, 1);
    int f = (lane == 0 || kprev != key) ? 1 : 0;
    seg_step<1>(v, f);
    seg_step<2>(v, f);
    seg_step<4>(v, f);
    seg_step<8>(v, f);
    seg_bcast<0x142, 0xa>(v, f);    // rows 1, 3 take the ends of rows 0, 2
    seg_bcast<0x143, 0xc>(v, f);    // rows 2, 3 take lane 31 (rows 0-1 combined)
    return lane == 63 || knext != key;
}

// ------------------------------------------------------- particle kernel
// Launch over every (species, tile) pair.  Tiles are contiguous slot
// ranges taken in slot order, block b first taking tile b (blocks are
// dispatched to the XCDs round-robin): the whole device then streams through
// one narrow stretch of every array at a time.  Cutting the tile range into
// eight contiguous parts, one per XCD, kept neighbouring tiles on one L2 but
// streamed eight distant stretches at once and made the pass 5 % slower.
// BLOCK threads (BLOCK / 64 waves) take a tile of kTile slots, ITEMS
// consecutive runs of 64 slots per wave; WINR is the deposit window's height.
// Each wave loads the next run's particles while it finishes the current
// one; the loads are issued after the current run's last dependent read (the
// field gather and the material lookup), because a wait for any later load
// would also wait for them.  The register budget is held at five waves per
// SIMD (four with split/merge folded in, which needs more registers than
// that budget leaves without spilling).  The fused pass is instantiated for
// several (BLOCK, WINR) pairs, which the autotuner chooses among; every
// other mode uses (256, 24).
// Waves per SIMD that the particle pass's LDS (deposit window and queue
// staging) leaves room for, at most the register budget's five, and a
// whole number of blocks (a 512-thread block puts two waves on each SIMD).
constexpr int pk_waves(int block, int winr)
{
    const int bytes = kWinZ * winr * 8 + block * kRecD * 8 + 512;
    const int per = block >= 256 ? block / 256 : 1;       // waves per SIMD of one block
    const int w = (160 * 1024 / bytes) * (block / 64) / 4;
    const int cap = 5 / per * per;
    return (w < cap ? w : cap) / per * per;
}
template <int MODE, int BLOCK = kBlock, int WINR = kWinR>
__global__ void __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu((MODE & M_REZONE) ? 4 : pk_waves(BLOCK, WINR))))
particle_kernel(const DevState *__restrict__ stp, DevCounters *__restrict__ ctr, int sel)
{
    constexpr int ITEMS = kTile / BLOCK;
    constexpr int kAnchorR = anchor_rows(WINR);
    const DevState &st = *stp;
    __shared__ double win[kWinZ * WINR];
    __shared__ unsigned long long dead_blk[IPM_NS_MAX];
    __shared__ double metal_blk[kMetalLds];     // this block's metal charge tallies
    // Entries of the tile in the collision queue so far (two counters so
    // the next tile never waits for the reset), and each wave's staging
    // buffer of up to 64 queue records, written out with full waves.
    __shared__ unsigned q_scount[2];
    __shared__ double q_wst[(MODE & M_MCC) ? (BLOCK / 64) * 64 * kRecD : 1];

    const int ns = st.ns;
    // The populations, the step and the seed are read through the scalar
    // cache (every wave needs the same values, and this pass never writes
    // them: it adds to n_new, n_dead, q_total and overflow only), so a
    // block's first loads of particles do not wait behind vector loads of
    // counters.
    const auto *cn = (const __attribute__((address_space(4))) unsigned long long *)ctr->n;
    long long T = 0;
    for (int s = 0; s < ns; ++s)
        if (sel < 0 || sel == s) T += ((long long)cn[s] + kTile - 1) / kTile;

    if (MODE & M_DEPOSIT)
        for (int q = threadIdx.x; q < kWinZ * WINR; q += BLOCK) win[q] = 0.0;
    if (threadIdx.x < IPM_NS_MAX) dead_blk[threadIdx.x] = 0;
    if (threadIdx.x < kMetalLds) metal_blk[threadIdx.x] = 0.0;
    if (threadIdx.x < 2) q_scount[threadIdx.x] = 0;
    __syncthreads();
    int qbuf = 0;

    const long long t1 = T, tfirst = blockIdx.x, tstride = gridDim.x;

    const uint32_t step = *(const __attribute__((address_space(4))) uint32_t *)&ctr->step;
    const unsigned long long seed = *(const __attribute__((address_space(4))) unsigned long long *)&ctr->seed;
    const int lane = threadIdx.x & 63;

    for (long long t = tfirst; t < t1; t += tstride) {
        // species owning tile t: walk the (at most 8) per-species tile counts
        int s = 0;
        long long tbeg = 0;
        for (int q = 0; q < ns; ++q) {
            const long long nt = (sel < 0 || sel == q) ? ((long long)cn[q] + kTile - 1) / kTile : 0;
            if (t < tbeg + nt) { s = q; break; }
            tbeg += nt;
        }
        const long long n_s = (long long)cn[s];
        const long long base = (t - tbeg) * kTile;
        double *const *F = st.f[s];
        const double charge = st.charge[s];
        const double efactor = st.efactor[s];
        double *__restrict__ raw = st.raw[s];

        // Deposit window of the tile: placed by the tile's first particle as
        // the previous pass left it (any placement is correct: cells outside
        // the window go straight to HBM), so no wave waits for a load here.
        const auto anc = gp(st.tile_anchor) + 2 * (st.q_cbase[s] + (t - tbeg));
        const int zw0 = (MODE & M_DEPOSIT) ? anc[0] : 0;
        const int rw0 = (MODE & M_DEPOSIT) ? anc[1] : 0;
        int na0 = zw0, na1 = rw0;   // next pass's window, from slot base after this pass

        // Wave w of the block takes ITEMS consecutive runs of 64 slots, so a
        // wave walks a contiguous, cell-sorted stretch of the tile.  The
        // particle streams are read and written once per pass: non-temporal
        // accesses keep them from evicting the field grids and the deposit
        // targets from L2.
        const int wv = threadIdx.x >> 6;
        auto load_item = [&](long long i, Part &q) {
            const unsigned bo = (unsigned)i << 3;
            q.Z = __builtin_nontemporal_load(at_off(F[0], bo));
            if (MODE & (M_PUSH | M_ABSORB | M_DEPOSIT)) {
                q.R = __builtin_nontemporal_load(at_off(F[1], bo));
                q.WQ = __builtin_nontemporal_load(at_off(F[7], bo));
            }
            if (MODE & (M_PUSH | M_ABSORB)) {
                q.Vz = __builtin_nontemporal_load(at_off(F[2], bo)); q.Vr = __builtin_nontemporal_load(at_off(F[3], bo));
                q.Vt = __builtin_nontemporal_load(at_off(F[4], bo));
                q.Az = __builtin_nontemporal_load(at_off(F[5], bo)); q.Ar = __builtin_nontemporal_load(at_off(F[6], bo));
            }
        };
        // Queue records staged by this wave (wave-uniform count).  A flush
        // reserves space in the tile's region with one LDS atomic and writes
        // the records with every lane storing 16 contiguous bytes.  The flush
        // at the end of the tile first drops the records whose collision is
        // surely a null event (one gas, a table with a screen): lane r tests
        // record r, and the kept ones are packed to the front.  Their
        // velocity already carries the VelRes round trip, so nothing else is
        // left of them but the event count.
        unsigned cw = 0;
        auto flush_wave = [&](bool screen) {
            if (screen) {
                const DevTable &tb = st.tab[s];
                double *wst = q_wst + wv * 64 * kRecD;
                const bool has = (unsigned)lane < cw;
                double rv[kRecD];
                bool keep = has;
                if (has) {
#pragma unroll
                    for (int f = 0; f < kRecD; ++f) rv[f] = wst[lane * kRecD + f];
                    keep = !surely_null(tb, rv[2] * st.vfactor, rv[3] * st.vfactor, rv[4] * st.vfactor, rv[7]);
                }
                const unsigned long long km = __ballot(keep);
                const unsigned nkeep = (unsigned)__popcll(km);
                if (st.count_events && lane == 0 && cw > nkeep)
                    global_fetch_add(&st.mcc_counts[(size_t)s * 128], (unsigned long long)(cw - nkeep));
                if (keep) {
                    const unsigned r = (unsigned)__popcll(km & ((1ull << lane) - 1ull));
#pragma unroll
                    for (int f = 0; f < kRecD; ++f) wst[r * kRecD + f] = rv[f];
                }
                cw = nkeep;
                if (!cw) return;
            }
            unsigned ob = 0;
            if (lane == 0) ob = atomicAdd(&q_scount[qbuf], cw);
            ob = __shfl(ob, 0);
            const long long R = st.q_tile_cap[s];
            double *qt = st.q_rec + (size_t)(st.q_region[s] + (t - tbeg) * R) * kRecD;
            const double *wst = q_wst + wv * 64 * kRecD;
            for (unsigned j = lane; j < cw * (unsigned)kRecParts; j += 64u) {
                const unsigned rec = j / (unsigned)kRecParts, part = j - rec * (unsigned)kRecParts;
                if ((long long)(ob + rec) < R) {
                    const rec2 v = {wst[rec * kRecD + part * 2], wst[rec * kRecD + part * 2 + 1]};
                    *(__attribute__((address_space(1))) rec2 *)(qt + (size_t)(ob + rec) * kRecD + part * 2) = v;
                } else if (part == 0u) {
                    atomicOr(&ctr->overflow, 4);
                }
            }
            cw = 0;
        };
        Part pn;   // the wave's next run of particles, loaded ahead
        {
            const long long i0 = base + (long long)(wv * ITEMS) * 64 + lane;
            if (i0 < n_s) load_item(i0, pn);
        }
#pragma unroll 1
        for (int k = 0; k < ITEMS; ++k) {
            const long long i = base + (long long)(wv * ITEMS + k) * 64 + lane;
            const bool valid = i < n_s;
            Part p = pn;
            bool live = false, deleted = false;
            if (valid) live = !is_dead(p.Z);
            if (live) {
                // Vacuum runs of four cell rows around the first live lane's
                // row before the push, read as scalars while the push's
                // gather is in flight (scalar loads do not wait on vector
                // loads); a particle moves less than a cell per step.
                RowRuns rr;
                if (MODE & M_ABSORB) {
                    const int row = min(max((int)ceil(p.R), 1), st.nr - 1) - 1;
                    rr.r0 = max(__builtin_amdgcn_readfirstlane(row) - 1, 0);
                    rr.n = 4;
                    const auto *rv = (const __attribute__((address_space(4))) int *)st.row_vac + 2 * rr.r0;
                    rr.v0 = int2{rv[0], rv[1]};
                    rr.v1 = int2{rv[2], rv[3]};
                    rr.v2 = int2{rv[4], rv[5]};
                    rr.v3 = int2{rv[6], rv[7]};
                }
                if (MODE & M_PUSH) push_one(st, efactor, p);
                if (MODE & M_ABSORB) deleted = absorb_one(st, charge, p, rr, metal_blk);
            }
            // every dependent read of this run is done: load the next one
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 < ITEMS && i + 64 < n_s) load_item(i + 64, pn);
            __builtin_amdgcn_sched_barrier(0);
            const bool alive = live && !deleted;
            if ((MODE & M_DEPOSIT) && threadIdx.x == 0 && k == 0 && alive && p.Z > 0.0 && p.R >= 0.0 &&
                p.Z < 1e9 && p.R < 1e9) {
                na0 = (int)ceil(p.Z) - kAnchorZ;
                na1 = (int)ceil(p.R) - kAnchorR;
            }

            // Null-collision selection (SelectParticleVW, R/src/mc/MCCPublic.F90:123-155):
            // draw 0 of the (slot, step, species|gas<<8) stream against
            // CollisionRatio.  Selected particles go to the collision queue
            // with their post-push state; collide_kernel finishes them.
            // Draw 1 (the other half of the same Philox block) is the
            // uniform of the reaction choice; it rides in the queue record.
            bool selected = false;
            double usel = 0.0;
            if ((MODE & M_MCC) && alive) {
                if (!(MODE & (M_PUSH | M_ABSORB))) {
                    if (!(MODE & M_DEPOSIT)) p.R = gp(F[1])[i];
                    p.Vz = gp(F[2])[i]; p.Vr = gp(F[3])[i]; p.Vt = gp(F[4])[i];
                }
                const StateKey sk = state_key(p.Z, p.R, p.Vz, p.Vr, p.Vt);
                for (int g = 0; g < st.ng; ++g) {
                    const DevTable &tb = st.tab[s * st.ng + g];
                    if (tb.model == 0) continue;
                    Rng rg = make_rng(sk, seed, step, (uint32_t)s | ((uint32_t)g << 8));
                    if (rg.next() < tb.ratio) {
                        selected = true;
                        usel = rg.next();
                    }
                }
            }
            if (MODE & M_DEPOSIT) {
                // P2CES, R/src/pic/field/FieldOne.f90:151-166.  Contributions
                // of a run of particles in one cell are summed across the wave
                // first; the run's last lane adds them to the LDS window (or to
                // HBM when the cell lies outside it).
                int key = -1 - lane;
                double w[4] = {0.0, 0.0, 0.0, 0.0};
                int NzL = 1, NrL = 1;
                if (alive) {
                    NzL = (int)ceil(p.Z);
                    double ZL = (double)NzL - p.Z, ZU = 1.0 - ZL;
                    NrL = (int)ceil(p.R);
                    double RL = (double)NrL - p.R, RU = 1.0 - RL;
                    NzL = min(max(NzL, 1), st.nz - 1);
                    const bool row0 = NrL < 1;            // R == 0 exactly: weight RL is 0
                    NrL = min(max(NrL, 1), st.nr - 1);
                    w[0] = ZL * RL * p.WQ; w[1] = ZL * RU * p.WQ;
                    w[2] = ZU * RL * p.WQ; w[3] = ZU * RU * p.WQ;
                    if (row0) {   // the reference's row-0 terms carry weight 0; row 1 takes the RU terms
                        w[0] = w[1]; w[2] = w[3]; w[1] = 0.0; w[3] = 0.0;
                    }
                    key = (NzL - 1) + (NrL - 1) * (st.nz - 1);
                }
                const bool run_end = wave_segmented_sum(key, w);
                if (run_end && key >= 0) {
                    const int wz = NzL - zw0, wr = NrL - rw0;
                    if (wz >= 0 && wz + 1 < kWinZ && wr >= 0 && wr + 1 < WINR) {
                        double *wp = win + wz + wr * kWinZ;
                        lds_add(wp, w[0]);
                        lds_add(wp + kWinZ, w[1]);
                        lds_add(wp + 1, w[2]);
                        lds_add(wp + 1 + kWinZ, w[3]);
                    } else {
                        double *g0 = raw + (size_t)(NzL - 1) + (size_t)(NrL - 1) * st.nz;
                        global_add(g0, w[0]);
                        global_add(g0 + st.nz, w[1]);
                        global_add(g0 + 1, w[2]);
                        global_add(g0 + 1 + st.nz, w[3]);
                    }
                }
            }

            // Split/merge of the species' particles (particle_split_merge_control,
            // R/src/base/Rezoning.F90:17-92) where OneStep runs it: after the
            // deposit (which used the weight before the split) and before the
            // collision pass (which sees the parent with its reduced weight and
            // every copy).  The same rules as split_merge_kernel; each copy is
            // written here and drawn for selection on its own stream.
            unsigned ncopy = 0;            // copies of this lane's particle (split)
            unsigned long long cbase = 0;  // their first slot
            double cref = 0.0;             // their weight
            bool wq_changed = false;
            if constexpr ((MODE & M_REZONE) != 0) {
                const int rzm = st.rz_mode[s];
                if (alive && rzm) {
                    const double ref = st.rz_weight[s] * ((1.0 - st.rz_alpha[s]) * p.R / st.rz_gridr[s] + st.rz_alpha[s]);
                    if (ref > 0.0) {
                        double wq = p.WQ;
                        if ((rzm & 1) && wq / ref >= 2.0) {
                            const double ratio = wq / ref;
                            const long long capS = st.cap[s];
                            if (!(ratio - 1.0 <= (double)(capS - n_s))) {
                                atomicOr(&ctr->overflow, 1);
                            } else {
                                const unsigned long long mcp = (unsigned long long)(long long)ratio - 1ull;
                                const unsigned long long k0 = atomicAdd(&ctr->n_new[s], mcp);
                                const unsigned long long b0 = (unsigned long long)n_s + k0;
                                if (b0 + mcp > (unsigned long long)capS) {
                                    atomicOr(&ctr->overflow, 1);
                                    unsigned long long nd = 0;
                                    for (unsigned long long q = b0; q < (unsigned long long)capS && q < b0 + mcp; ++q, ++nd)
                                        gp(F[0])[q] = kDead;
                                    if (nd) atomicAdd(&ctr->n_dead[s], nd);
                                } else {
                                    for (unsigned long long j = 1; j <= mcp; ++j) wq = wq - ref;
                                    ncopy = (unsigned)mcp;
                                    cbase = b0;
                                    cref = ref;
                                }
                            }
                        }
                        if ((rzm & 2) && wq < st.rz_limit[s] * ref) {
                            const StateKey mk = state_key(p.Z, p.R, p.Vz, p.Vr, p.Vt);
                            Rng g = make_rng(mk, seed, step, kMergeTag | (uint32_t)s);
                            if (g.next() < wq / ref) {
                                wq = ref;
                            } else {
                                deleted = true;
                                selected = false;
                            }
                        }
                        if (__double_as_longlong(wq) != __double_as_longlong(p.WQ)) {
                            p.WQ = wq;
                            wq_changed = true;
                        }
                    }
                }
            }

            if (MODE & M_MCC) {
                // stage the wave's selections as queue records in LDS
                const unsigned long long m = __ballot(selected);
                const unsigned nsel = (unsigned)__popcll(m);
                if (nsel) {
                    if (cw + nsel > 64u) flush_wave(false);
                    if (selected) {
                        if (!(MODE & (M_PUSH | M_ABSORB | M_DEPOSIT))) p.WQ = gp(F[7])[i];
                        const unsigned r = cw + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                        double *d = q_wst + wv * 64 * kRecD + r * kRecD;
                        d[0] = p.Z; d[1] = p.R; d[2] = p.Vz; d[3] = p.Vr;
                        d[4] = p.Vt; d[5] = p.WQ;
                        d[6] = __longlong_as_double((long long)(i - base));
                        d[7] = usel;
                    }
                    cw += nsel;
                }
            }
            if constexpr ((MODE & M_REZONE) != 0) {
                // Copy j (1-based) of a split parent: the parent's state with
                // the low bits of Vt's pattern XOR j and weight ref, written at
                // its reserved slot; drawn for selection like any particle, its
                // queue record names the slot relative to this tile.
                unsigned jmax = ncopy;
                for (int o = 32; o > 0; o >>= 1) jmax = max(jmax, (unsigned)__shfl_xor((int)jmax, o));
                const unsigned long long vtb = (unsigned long long)__double_as_longlong(p.Vt);
                for (unsigned j = 1; j <= jmax; ++j) {
                    const bool mine = j <= ncopy;
                    bool csel = false;
                    double ucp = 0.0;
                    Part cp = p;
                    unsigned long long q = 0;
                    if (mine) {
                        q = cbase + j - 1;
                        cp.Vt = __longlong_as_double((long long)(vtb ^ (unsigned long long)j));
                        cp.WQ = cref;
                        if (MODE & M_MCC) {
                            const StateKey ck = state_key(cp.Z, cp.R, cp.Vz, cp.Vr, cp.Vt);
                            for (int g = 0; g < st.ng; ++g) {
                                const DevTable &tb = st.tab[s * st.ng + g];
                                if (tb.model == 0) continue;
                                Rng rg = make_rng(ck, seed, step, (uint32_t)s | ((uint32_t)g << 8));
                                if (rg.next() < tb.ratio) {
                                    csel = true;
                                    ucp = rg.next();
                                }
                            }
                        }
                        // a selected copy's stored velocity carries the VelRes
                        // round trip, as a selected particle's does (one gas)
                        double vz = cp.Vz, vr = cp.Vr, vt = cp.Vt;
                        if ((MODE & M_MCC) && csel && st.ng == 1) {
                            vz = vz * st.vfactor * st.inv_vfactor;
                            vr = vr * st.vfactor * st.inv_vfactor;
                            vt = vt * st.vfactor * st.inv_vfactor;
                        }
                        F[0][q] = cp.Z; F[1][q] = cp.R; F[2][q] = vz; F[3][q] = vr;
                        F[4][q] = vt; F[5][q] = cp.Az; F[6][q] = cp.Ar; F[7][q] = cp.WQ;
                    }
                    if (MODE & M_MCC) {
                        const unsigned long long m = __ballot(csel);
                        const unsigned nsel = (unsigned)__popcll(m);
                        if (nsel) {
                            if (cw + nsel > 64u) flush_wave(false);
                            if (csel) {
                                const unsigned r = cw + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                                double *d = q_wst + wv * 64 * kRecD + r * kRecD;
                                d[0] = cp.Z; d[1] = cp.R; d[2] = cp.Vz; d[3] = cp.Vr;
                                d[4] = cp.Vt; d[5] = cp.WQ;
                                d[6] = __longlong_as_double((long long)q - base);
                                d[7] = ucp;
                            }
                            cw += nsel;
                        }
                    }
                }
            }

            if (MODE & (M_ABSORB | M_REZONE)) {
                const unsigned long long m = __ballot(deleted);
                if (m && lane == (int)(__ffsll((long long)m) - 1))
                    atomicAdd(&dead_blk[s], (unsigned long long)__popcll(m));
            }
            if (deleted) {
                gp(F[0])[i] = kDead;
            } else if (alive) {
                // With one gas, a selected particle's VelRes(VFactor) and
                // VelRes(1/VFactor) (R/src/mc/MCCPublic.F90:62, 91) are all a
                // null event leaves, so they are applied here, in the
                // streaming store, and the collision pass writes back only
                // real reactions (the queue record keeps the unrounded state).
                const bool rt = (MODE & M_MCC) && selected && st.ng == 1;
                double vz = p.Vz, vr = p.Vr, vt = p.Vt;
                if (rt) {
                    vz = vz * st.vfactor * st.inv_vfactor;
                    vr = vr * st.vfactor * st.inv_vfactor;
                    vt = vt * st.vfactor * st.inv_vfactor;
                }
                if (MODE & M_PUSH) {
                    const unsigned bo = (unsigned)i << 3;
                    __builtin_nontemporal_store(p.Z, at_off(F[0], bo)); __builtin_nontemporal_store(p.R, at_off(F[1], bo));
                    __builtin_nontemporal_store(vz, at_off(F[2], bo)); __builtin_nontemporal_store(vr, at_off(F[3], bo));
                    __builtin_nontemporal_store(vt, at_off(F[4], bo)); __builtin_nontemporal_store(p.Az, at_off(F[5], bo));
                    __builtin_nontemporal_store(p.Ar, at_off(F[6], bo));
                    if ((MODE & M_REZONE) && wq_changed) __builtin_nontemporal_store(p.WQ, at_off(F[7], bo));
                } else {
                    if (MODE & M_ABSORB) {
                        gp(F[1])[i] = p.R; gp(F[3])[i] = vr; gp(F[6])[i] = p.Ar;
                    }
                    if ((MODE & M_MCC) && rt) {
                        gp(F[2])[i] = vz; gp(F[3])[i] = vr; gp(F[4])[i] = vt;
                    }
                }
            }
        }

        if (MODE & M_MCC) {
            if (cw) flush_wave(st.null_screen && st.ng == 1 && st.tab[s].nb_tab != nullptr);
        }
        if (MODE & M_DEPOSIT) {
            __syncthreads();
            for (int q = threadIdx.x; q < kWinZ * WINR; q += BLOCK) {
                const double v = win[q];
                if (v != 0.0) {
                    const int iz = zw0 + (q % kWinZ), ir = rw0 + (q / kWinZ);
                    global_add(raw + (size_t)(iz - 1) + (size_t)(ir - 1) * st.nz, v);
                    win[q] = 0.0;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) { anc[0] = na0; anc[1] = na1; }   // every wave has read anc by now
        }
        if (MODE & M_MCC) {
            // publish the tile's entry count (the deposit flush above already
            // ordered every wave's LDS counter update)
            if (!(MODE & M_DEPOSIT)) __syncthreads();
            if (threadIdx.x == 0) {
                const unsigned n = q_scount[qbuf];
                const long long R = st.q_tile_cap[s];
                const unsigned nq = (long long)n < R ? n : (unsigned)R;
                gp(st.q_tile_count)[st.q_cbase[s] + (t - tbeg)] = nq;
                if (nq) atomicAdd(&ctr->q_total[s], (unsigned long long)nq);
                q_scount[qbuf] = 0;     // reused two tiles later, after a barrier
            }
            qbuf ^= 1;
        }
    }

    if (MODE & M_ABSORB) {
        // The last tile's barrier (deposit flush or queue append) already
        // ordered every wave's dead_blk and metal_blk updates before these reads.
        if (!(MODE & (M_DEPOSIT | M_MCC))) __syncthreads();
        if (threadIdx.x < (unsigned)ns && dead_blk[threadIdx.x])
            atomicAdd(&ctr->n_dead[threadIdx.x], dead_blk[threadIdx.x]);
        if (threadIdx.x < (unsigned)min(st.nmetal, kMetalLds) && metal_blk[threadIdx.x] != 0.0)
            global_add(&st.metal_charge[threadIdx.x], metal_blk[threadIdx.x]);
    }

}

// ------------------------------------------------------- collision pass
// Three kernels finish the queued particles, with no atomic that many
// threads contend for (device-wide atomics on one address serialize, and
// the reference's append of products, MCCAddNewParticle,
// R/src/mc/MCCPublic.F90:209-226, would otherwise put one on every product).
//   collide_kernel: for each queued particle and each gas in order, the
//     selection draw (repeated with several gases; with one the particle
//     pass made it), VelRes(VFactor), Updater, the reaction choice and the
//     reaction of MCCPublic.F90:61-97, VelRes back; the changed fields are
//     written back and products staged in LDS, then in the block's own slice
//     of a global stage.  With one gas the particle pass has already dropped
//     the surely null events (and stored every selected particle's VelRes
//     round trip), so nearly every queued particle reacts and wavefronts run
//     the reaction code full; with several gases this kernel skips the
//     target sampling of surely null events itself.
//   collide_append_kernel: block b appends its staged products after the
//     products of blocks 0..b-1 of the same species.
//   collide_fold_kernel: folds the step's products into n[] and advances the
//     random stream's step key.
// collide_kernel splits its grid over the species in proportion to their
// expected work (every species with live tiles gets a block); a block takes
// a contiguous range of its species' tiles, scans their entry counts in LDS
// and processes the entries flat.  All three use fixed grids, so they can be
// captured.
constexpr int kChunk = 1024;   // tile counts scanned per LDS pass

__global__ void __launch_bounds__(kBlock, 2)
collide_kernel(const DevState *__restrict__ stp, DevCounters *__restrict__ ctr)
{
    const DevState &st = *stp;
    __shared__ unsigned cpre[kChunk + 1];
    __shared__ unsigned wtot[kBlock / 64];
    __shared__ double prec[kProdStage * 8];
    __shared__ unsigned pmeta[kProdStage];
    __shared__ unsigned pcount, pspc[IPM_NS_MAX];
    const int ns = st.ns;
    const unsigned G = gridDim.x;
    if (threadIdx.x == 0) pcount = 0;
    const ProdSink sink = {(lds_ptr<double>)prec, (lds_ptr<unsigned>)pmeta, (lds_ptr<unsigned>)&pcount,
                           (lds_ptr<unsigned>)pspc};
    if (threadIdx.x < IPM_NS_MAX) pspc[threadIdx.x] = 0;

    // species of this block and its tile range: one block for every species
    // with MCC tables and live tiles, the others in proportion to the
    // entries the particle pass queued for each species (after the null
    // screen the electrons keep a much smaller share of their selections
    // than ions, whose selections are nearly all real)
    int sb = -1;
    long long t0 = 0, t1 = 0;
    {
        double W = 0.0;
        int A = 0;
        for (int s = 0; s < ns; ++s) {
            const long long nt = ((long long)ctr->n[s] + kTile - 1) / kTile;
            if (nt > 0 && st.q_weight[s] > 0.0) { W += (double)ctr->q_total[s] + 1.0; ++A; }
        }
        // every species with work needs a block of its own: a grid smaller
        // than that would leave queued selections uncollided
        if (A > (int)G && blockIdx.x == 0 && threadIdx.x == 0) global_or(&ctr->overflow, 4);
        if (A > 0 && G >= (unsigned)A) {
            const double Grem = (double)(G - A);
            double Wlt = 0.0;
            long long B = 0;
            for (int s = 0; s < ns; ++s) {
                const long long nt = ((long long)ctr->n[s] + kTile - 1) / kTile;
                if (!(nt > 0 && st.q_weight[s] > 0.0)) continue;
                const double Wle = Wlt + (double)ctr->q_total[s] + 1.0;
                const long long nb = 1 + (long long)floor(Grem * Wle / W) - (long long)floor(Grem * Wlt / W);
                if ((long long)blockIdx.x >= B && (long long)blockIdx.x < B + nb) {
                    const long long lb = (long long)blockIdx.x - B;
                    sb = s;
                    t0 = nt * lb / nb;
                    t1 = nt * (lb + 1) / nb;
                }
                B += nb;
                Wlt = Wle;
            }
        }
    }
    const uint32_t step = ctr->step;
    const unsigned long long seed = ctr->seed;
    if (sb >= 0) {
        const int s = sb;
        const long long R = st.q_tile_cap[s];
        const unsigned *cnt = st.q_tile_count + st.q_cbase[s];
        double *const *F = st.f[s];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (long long tc = t0; tc < t1; tc += kChunk) {
            const int m = (int)min((long long)kChunk, t1 - tc);
            // exclusive scan of the chunk's tile counts
            constexpr int per = kChunk / kBlock;
            unsigned c[per];
            unsigned tsum = 0;
            for (int j = 0; j < per; ++j) {
                const int tt = threadIdx.x * per + j;
                c[j] = tt < m ? gp(cnt)[tc + tt] : 0u;
                tsum += c[j];
            }
            unsigned x = tsum;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned y = __shfl_up(x, o);
                if (lane >= o) x += y;
            }
            if (lane == 63) wtot[w] = x;
            __syncthreads();
            unsigned before = 0;
            for (int k = 0; k < w; ++k) before += wtot[k];
            unsigned run = before + x - tsum;
            for (int j = 0; j < per; ++j) {
                cpre[threadIdx.x * per + j] = run;
                run += c[j];
            }
            if (threadIdx.x == kBlock - 1) cpre[kChunk] = run;
            __syncthreads();
            const unsigned total = cpre[kChunk];
            for (unsigned e = threadIdx.x; e < total; e += kBlock) {
                // tile of entry e: last j with cpre[j] <= e
                int lo = 0, hi = m;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (cpre[mid] <= e) lo = mid;
                    else hi = mid;
                }
                const long long lt = tc + lo;
                const long long q = st.q_region[s] + lt * R + (e - cpre[lo]);
                Part p;
                double urec;
                const long long i = lt * kTile + load_rec(st.q_rec + (size_t)q * kRecD, p, urec);
                const StateKey sk = state_key(p.Z, p.R, p.Vz, p.Vr, p.Vt);
                // the queue record does not carry Az and Ar
                p.Az = gp(F[5])[i];
                p.Ar = gp(F[6])[i];
                const Part p0 = p;
                for (int g = 0; g < st.ng; ++g) {
                    const DevTable &tb = st.tab[s * st.ng + g];
                    if (tb.model == 0) continue;
                    Rng rg = make_rng(sk, seed, step, (uint32_t)s | ((uint32_t)g << 8));
                    double u;
                    if (st.ng == 1) {
                        // the particle pass made draws 0 (selection) and 1
                        rg.seek(2);
                        u = urec;
                    } else {
                        if (!(rg.next() < tb.ratio)) continue;
                        u = rg.next();
                        if (st.null_screen && tb.nb_tab &&
                            surely_null(tb, p.Vz * st.vfactor, p.Vr * st.vfactor, p.Vt * st.vfactor, u)) {
                            // a null event: the count and VelRes both ways
                            if (st.count_events)
                                global_fetch_add(&st.mcc_counts[((size_t)s * st.ng + g) * 128], 1ull);
                            p.Vz = p.Vz * st.vfactor * st.inv_vfactor;
                            p.Vr = p.Vr * st.vfactor * st.inv_vfactor;
                            p.Vt = p.Vt * st.vfactor * st.inv_vfactor;
                            continue;
                        }
                    }
                    Mc mc;
                    const int idx = mcc_begin(st, s, tb, p, mc, rg, u);
                    mcc_finish(st, ctr, s, g, tb, mc, rg, idx, p, &sink);
                }
                // write back what the collisions changed: with one gas the
                // particle pass stored the VelRes round trip, all a null event
                // leaves; Z, R and WQ never change.  Each field is a separate
                // cache line of a separate array, stored non-temporally so the
                // scattered lines do not crowd L2.
                double vz0 = p0.Vz, vr0 = p0.Vr, vt0 = p0.Vt;
                if (st.ng == 1) {
                    vz0 = vz0 * st.vfactor * st.inv_vfactor;
                    vr0 = vr0 * st.vfactor * st.inv_vfactor;
                    vt0 = vt0 * st.vfactor * st.inv_vfactor;
                }
                if (__double_as_longlong(p.Vz) != __double_as_longlong(vz0) ||
                    __double_as_longlong(p.Vr) != __double_as_longlong(vr0) ||
                    __double_as_longlong(p.Vt) != __double_as_longlong(vt0)) {
                    __builtin_nontemporal_store(p.Vz, gp(F[2]) + i);
                    __builtin_nontemporal_store(p.Vr, gp(F[3]) + i);
                    __builtin_nontemporal_store(p.Vt, gp(F[4]) + i);
                }
                if (__double_as_longlong(p.Az) != __double_as_longlong(p0.Az) ||
                    __double_as_longlong(p.Ar) != __double_as_longlong(p0.Ar)) {
                    __builtin_nontemporal_store(p.Az, gp(F[5]) + i);
                    __builtin_nontemporal_store(p.Ar, gp(F[6]) + i);
                }
            }
            __syncthreads();
        }
    }
    // the block's products to its slice of the stage
    __syncthreads();
    const unsigned np = min(pcount, (unsigned)kProdStage);
    double *gs = st.p_stage + (size_t)blockIdx.x * kProdStage * 8;
    for (unsigned k = threadIdx.x; k < np * 8u; k += kBlock) gp(gs)[k] = prec[k];
    for (unsigned k = threadIdx.x; k < np; k += kBlock) gp(st.p_meta)[(size_t)blockIdx.x * kProdStage + k] = pmeta[k];
    if (threadIdx.x < IPM_NS_MAX) gp(st.p_cnt)[(size_t)blockIdx.x * IPM_NS_MAX + threadIdx.x] = pspc[threadIdx.x];
}

// Block b appends its staged products of species d after n[d], the
// directly reserved products (n_new[d]) and the staged products of blocks
// 0..b-1.
__global__ void __launch_bounds__(kBlock)
collide_append_kernel(const DevState *__restrict__ stp, DevCounters *ctr)
{
    const DevState &st = *stp;
    __shared__ unsigned long long off[IPM_NS_MAX];
    __shared__ unsigned long long part[kBlock / 64][IPM_NS_MAX];
    const int ns = st.ns;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // the collision pass has read the pass's queue totals: clear them for the next
    if (blockIdx.x == 0 && threadIdx.x < IPM_NS_MAX) ctr->q_total[threadIdx.x] = 0;
    unsigned long long acc[IPM_NS_MAX];
    for (int d = 0; d < IPM_NS_MAX; ++d) acc[d] = 0;
    for (unsigned b = threadIdx.x; b < blockIdx.x; b += kBlock)
        for (int d = 0; d < ns; ++d) acc[d] += gp(st.p_cnt)[(size_t)b * IPM_NS_MAX + d];
    for (int d = 0; d < ns; ++d) {
        unsigned long long x = acc[d];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
        if (lane == 0) part[w][d] = x;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)ns) {
        unsigned long long t = 0;
        for (int k = 0; k < kBlock / 64; ++k) t += part[k][threadIdx.x];
        off[threadIdx.x] = gp(ctr->n)[threadIdx.x] + gp(ctr->n_new)[threadIdx.x] + t;
    }
    __syncthreads();
    unsigned np = 0;
    for (int d = 0; d < ns; ++d) np += gp(st.p_cnt)[(size_t)blockIdx.x * IPM_NS_MAX + d];
    const auto *gs = gp(st.p_stage) + (size_t)blockIdx.x * kProdStage * 8;
    for (unsigned k = threadIdx.x; k < np; k += kBlock) {
        const unsigned meta = gp(st.p_meta)[(size_t)blockIdx.x * kProdStage + k];
        const int d = (int)(meta & 0xffu);
        const unsigned long long pos = off[d] + (meta >> 8);
        if (pos >= (unsigned long long)st.cap[d]) {
            global_or(&ctr->overflow, 1);
            continue;
        }
        double *const *F = st.f[d];
#pragma unroll
        for (int f = 0; f < 8; ++f) gp(F[f])[pos] = gs[(size_t)k * 8 + f];
    }
}

// Folds the step's products into n[] (capped at the capacity; the append
// pass flags what did not fit) and advances the random stream's step key.
// n[d] += the products of the collision pass (staged per block, plus the
// directly reserved ones), capped at the capacity; step += 1.  One block of
// kBlock threads.
__device__ __forceinline__ void fold_counts(const DevState &st, DevCounters *__restrict__ ctr)
{
    __shared__ unsigned long long tot[kBlock / 64][IPM_NS_MAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = 0; d < st.ns; ++d) {
        unsigned long long x = 0;
        for (int b = threadIdx.x; b < st.c_grid; b += blockDim.x) x += st.p_cnt[(size_t)b * IPM_NS_MAX + d];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
        if (lane == 0) tot[w][d] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int d = 0; d < st.ns; ++d) {
            unsigned long long add = ctr->n_new[d];
            for (int k = 0; k < (int)(blockDim.x / 64); ++k) add += tot[k][d];
            const unsigned long long n0 = ctr->n[d], cap = (unsigned long long)st.cap[d];
            ctr->n[d] = (n0 + add > cap) ? cap : n0 + add;
            ctr->n_new[d] = 0;
        }
        ctr->step = ctr->step + 1;
    }
}

__global__ void __launch_bounds__(kBlock) collide_fold_kernel(const DevState *__restrict__ stp,
                                                              DevCounters *__restrict__ ctr)
{
    fold_counts(*stp, ctr);
}

// ------------------------------------------------------- grid epilogue
// P2CES epilogue (R/src/pic/field/FieldOne.f90:168-175) for species in
// [s0, s1) and, if sum != 0, FieldSource%Sum (R/src/pic/field/FieldSource.f90:77-87).
// Raw deposits are cleared here so that no memset enters a captured graph.
// With fold != 0 (the fused step, after the collision pass) block 0 also
// folds the collision products into the counts (fold_counts).
__global__ void __launch_bounds__(kBlock)
grid_kernel(const DevState *__restrict__ stp, DevCounters *__restrict__ ctr, const double *__restrict__ inv_vol,
            const double *__restrict__ chifac, double *__restrict__ rho_one, double *__restrict__ chi_one,
            double *__restrict__ rho_tot, double *__restrict__ chi_tot, int s0, int s1, int sum, int fold)
{
    const DevState &st = *stp;
    if (fold && blockIdx.x == 0) fold_counts(st, ctr);
    const int nz = st.nz, nr = st.nr;
    const size_t nn = (size_t)nz * nr;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ng = (size_t)(nz + 2) * (nr + 2);
    if (tid < nn) {
        for (int s = s0; s < s1; ++s) {
            double v = st.raw[s][tid] * inv_vol[tid];
            v = v * st.charge[s];
            rho_one[(size_t)s * nn + tid] = v;
            chi_one[(size_t)s * nn + tid] = v * chifac[s];
            st.raw[s][tid] = 0.0;
        }
        if (sum) {
            double r = 0.0, c = 0.0;
            for (int s = 0; s < st.ns; ++s) {
                r = r + rho_one[(size_t)s * nn + tid];
                c = c + chi_one[(size_t)s * nn + tid];
            }
            rho_tot[tid] = r;
            const size_t iz = tid % nz, ir = tid / nz;
            chi_tot[(iz + 1) + (ir + 1) * (size_t)(nz + 2)] = c;
        }
    }
    if (sum && tid < ng) {
        const size_t iz = tid % (nz + 2), ir = tid / (nz + 2);
        if (iz == 0 || ir == 0 || iz == (size_t)nz + 1 || ir == (size_t)nr + 1) chi_tot[tid] = 0.0;
    }
}

// ------------------------------------------------------- moment diagnostics
// WeightingParticleField for one species, R/src/diag/diagParticleField.F90:196-257:
// node sums, at the P2CES node weights, of WQ, WQ*(Vr^2+Vt^2+Vz^2), 1 (the
// macro-particle count), WQ*Vz and WQ*Vr.  With implicit_j the currents use
// the velocity of a copy pushed once more in the present field
// (MoveParticleOneElectrostaticImplicit, R/src/pic/particle/ParticleBundle.f90:782-858),
// as the reference does for its implicit scheme (diagParticleField.F90:225-236).
__global__ void __launch_bounds__(kBlock)
moments_kernel(const DevState *__restrict__ stp, const DevCounters *__restrict__ ctr, int s, int implicit_j,
               double *__restrict__ acc)
{
    const DevState &st = *stp;
    const long long n = (long long)ctr->n[s];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if ((long long)blockIdx.x * blockDim.x >= n) return;          // whole block past the live slots
    const int lane = threadIdx.x & 63;
    double *const *F = st.f[s];
    // Every lane takes part in the wave sums: lanes without a particle carry
    // zero weights and a key of their own.
    int key = -1 - lane;
    double wr[4] = {0, 0, 0, 0}, we[4] = {0, 0, 0, 0}, wm[4] = {0, 0, 0, 0}, wz[4] = {0, 0, 0, 0},
           wrr[4] = {0, 0, 0, 0};
    int NzL = 1, NrL = 1;
    Part p;
    p.Z = i < n ? gp(F[0])[i] : kDead;
    if (!is_dead(p.Z)) {
        p.R = gp(F[1])[i]; p.Vz = gp(F[2])[i]; p.Vr = gp(F[3])[i]; p.Vt = gp(F[4])[i];
        p.Az = gp(F[5])[i]; p.Ar = gp(F[6])[i]; p.WQ = gp(F[7])[i];
        NzL = (int)ceil(p.Z);
        const double ZL = (double)NzL - p.Z, ZU = 1.0 - ZL;
        NrL = (int)ceil(p.R);
        const double RL = (double)NrL - p.R, RU = 1.0 - RL;
        NzL = min(max(NzL, 1), st.nz - 1);
        const bool row0 = NrL < 1;
        NrL = min(max(NrL, 1), st.nr - 1);
        const double energy = p.Vr * p.Vr + p.Vt * p.Vt + p.Vz * p.Vz;
        double vz = p.Vz, vr = p.Vr;
        if (implicit_j) {
            Part t = p;
            push_one(st, st.efactor[s], t);
            vz = t.Vz;
            vr = t.Vr;
        }
        double w[4] = {ZL * RL, ZL * RU, ZU * RL, ZU * RU};
        if (row0) { w[0] = w[1]; w[2] = w[3]; w[1] = 0.0; w[3] = 0.0; }
        for (int q = 0; q < 4; ++q) {
            const double ww = w[q] * p.WQ;
            wr[q] = ww; we[q] = ww * energy; wm[q] = w[q]; wz[q] = ww * vz; wrr[q] = ww * vr;
        }
        key = (NzL - 1) + (NrL - 1) * (st.nz - 1);
    }
    // runs of particles in one cell (cell-sorted storage) are summed across
    // the wave first; only the run's last lane adds to memory
    wave_segmented_sum(key, wr);
    wave_segmented_sum(key, we);
    wave_segmented_sum(key, wm);
    wave_segmented_sum(key, wz);
    const bool run_end = wave_segmented_sum(key, wrr);
    if (run_end && key >= 0) {
        const size_t nn = (size_t)st.nz * st.nr;
        const size_t k0 = (size_t)(NzL - 1) + (size_t)(NrL - 1) * st.nz;
        const size_t node[4] = {k0, k0 + st.nz, k0 + 1, k0 + 1 + st.nz};
        for (int q = 0; q < 4; ++q) {
            if (wm[q] == 0.0) continue;
            global_add(acc + node[q], wr[q]);
            global_add(acc + nn + node[q], we[q]);
            global_add(acc + 2 * nn + node[q], wm[q]);
            global_add(acc + 3 * nn + node[q], wz[q]);
            global_add(acc + 4 * nn + node[q], wrr[q]);
        }
    }
}

// Scaling of the sums, R/src/diag/diagParticleField.F90:300-318: densities
// times point_inv_volume, the energy density times 0.5 m VFactor^2 and
// divided by JtoeV and the density where that is positive (mean energy in
// eV), currents times point_inv_volume * Charge * VFactor.
__global__ void moments_scale_kernel(double *__restrict__ acc, const double *__restrict__ inv_vol, size_t nn,
                                     double energy_factor, double charge, double vfactor)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const double iv = inv_vol[k];
    const double rho = acc[k] * iv;
    double e = acc[nn + k] * iv;
    e = e * energy_factor;
    if (rho > 0.0) e = e / kJTOEV / rho;
    acc[k] = rho;
    acc[nn + k] = e;
    acc[3 * nn + k] = acc[3 * nn + k] * iv * charge * vfactor;
    acc[4 * nn + k] = acc[4 * nn + k] * iv * charge * vfactor;
}

// FT%SetE(FG), R/src/pic/field/FieldSource.f90:95-120, on a potential whose
// ghost layers are filled by linear extrapolation, the reference's default
// on every side of a one-rank domain (R/src/pic/domain/PICCom.F90:40-43,
// 516-578: ghost = 2*edge - next).  One thread per face of Ez then of Er.
__global__ void set_e_kernel(const double *__restrict__ phi, double *__restrict__ Ez, double *__restrict__ Er,
                             int nz, int nr, double dz)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nez = (long long)(nz + 1) * nr, ner = (long long)nz * (nr + 1);
    auto P = [&](int i, int j) -> double {       // potential with ghosts, i = 0..nz+1, j = 0..nr+1
        auto in = [&](int a, int b) { return phi[(size_t)(a - 1) + (size_t)(b - 1) * nz]; };
        if (i == 0) return 2.0 * in(1, j) - in(2, j);
        if (i == nz + 1) return 2.0 * in(nz, j) - in(nz - 1, j);
        if (j == 0) return 2.0 * in(i, 1) - in(i, 2);
        if (j == nr + 1) return 2.0 * in(i, nr) - in(i, nr - 1);
        return in(i, j);
    };
    if (t < nez) {
        const int i = (int)(t % (nz + 1)), j = (int)(t / (nz + 1)) + 1;   // Ez(i, j), i = 0..nz
        const double v = P(i, j) - P(i + 1, j);
        Ez[t] = v / dz;
    } else if (t < nez + ner) {
        const long long u = t - nez;
        const int i = (int)(u % nz) + 1, j = (int)(u / nz);             // Er(i, j), j = 0..nr
        const double v = P(i, j) - P(i, j + 1);
        Er[u] = v / dz;
    }
}

// FieldSolver%SetCoe (with_chi = 1) or %SetCoeWN (0), then %SetSrc and
// %SetBound (R/src/pic/field/FieldSolver.f90:136-301), one thread per node
// (i, j).  The five coefficients multiply phi(i-1, j), phi(i, j), phi(i+1, j),
// phi(i, j-1) and phi(i, j+1).  Each product and sum is evaluated in the
// reference's left-to-right order.  cell_epsilon is read at the Fortran
// linear address (a-1) + (b-1)(Nz-1) of its (Nz-1, Nr-1) array, as the
// reference reads it: on the axis next to the z = L boundary (i = Nz-1) CoeC
// names cell (Nz, 1), which that address places at (1, 2).  The host checks
// Nz, Nr >= 3, so every address stays inside the array.
__global__ void solver_system_kernel(const double *__restrict__ chi, const double *__restrict__ eps,
                                     const int *__restrict__ ptype, const double *__restrict__ rho,
                                     const double *__restrict__ sigma, const double *__restrict__ inv_vol,
                                     const double *__restrict__ bound, const int *__restrict__ mask, int with_chi,
                                     double scale, double *__restrict__ out, int nz, int nr)
{
    const long long nn = (long long)nz * nr;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int i = (int)(k % nz) + 1, j = (int)(k / nz) + 1;
    auto X = [&](int a, int b) { return chi[(size_t)a + (size_t)b * (nz + 2)]; };          // Chi(0:Nz+1, 0:Nr+1)
    auto EP = [&](int a, int b) { return eps[(size_t)(a - 1) + (size_t)(b - 1) * (nz - 1)]; };
    auto F = [&](int a, int b) { return with_chi ? 1.0 + 0.5 * (X(i, j) + X(a, b)) : 1.0; };
    double A = 0.0, B = 0.0, C = 0.0, D = 0.0, E = 0.0;
    if (ptype[k] > 0) {                       // metal node: phi is the electrode's
        B = 1.0;
    } else if (j == nr || i == 1 || i == nz) {
        // outer boundary: the solver takes these rows from the boundary values
    } else if (j == 1) {                      // axis
        A = F(i - 1, j) * 0.5 * (EP(i, j) + EP(i - 1, j));
        C = F(i + 1, j) * 0.5 * (EP(i, j) + EP(i + 1, j));
        E = 4.0 * F(i, j + 1) * 0.5 * (EP(i, j) + EP(i, j + 1));
        B = -(A + C + D + E);
    } else {
        const double jd = (double)j;
        A = F(i - 1, j) * 0.5 * (EP(i - 1, j - 1) + EP(i - 1, j));
        C = F(i + 1, j) * 0.5 * (EP(i, j - 1) + EP(i, j));
        D = F(i, j - 1) * 0.5 * (EP(i - 1, j - 1) + EP(i, j - 1)) * (jd - 1.5) / (jd - 1.0);
        E = F(i, j + 1) * 0.5 * (EP(i - 1, j) + EP(i, j)) * (jd - 0.5) / (jd - 1.0);
        B = -(A + C + D + E);
    }
    out[k] = A;
    out[nn + k] = B;
    out[2 * nn + k] = C;
    out[3 * nn + k] = D;
    out[4 * nn + k] = E;
    const double sg = sigma ? sigma[k] : 0.0;
    double src = scale * (rho[k] + sg * inv_vol[k]);
    if (mask && mask[k] != 0) src = bound[k];
    out[5 * nn + k] = src;
}

__global__ void clear_kernel(double *a, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// ------------------------------------------------------------ rezoning

// split_particle_weight followed by merge_particle_random,
// R/src/base/Rezoning.F90:53-92, one thread per slot.  The reference weight
// of a particle is CalRW, R/src/pic/particle/ParticleBundle.f90:393-402:
// Weight * ((1 - alpha) * R / grid_r + alpha).
//   split  (mode & 1): Ncr = int(WQ / ref); for Ncr > 1 the particle gives
//          Ncr - 1 copies of weight ref, subtracting ref once per copy.
//   merge  (mode & 2): a particle with WQ < merge_limit * ref survives with
//          weight ref when its uniform draw is below WQ / ref and is deleted
//          otherwise.  The copies of a split never qualify (ref < l * ref
//          implies a draw below 1: always kept at weight ref), so only the
//          particle itself is tested, after its split.
// Copies are appended after the n slots in use (folded by rezone_fold_kernel).
// Copy j (1-based) of a particle has the low bits of its Vt bit pattern
// XORed with j: copies then key different random streams in the collision
// and merge passes, where exact copies would draw alike forever.  A particle
// whose ref is not positive (alpha = 0 on the axis) is left alone.
__global__ void __launch_bounds__(kBlock)
split_merge_kernel(const DevState *__restrict__ stp, DevCounters *__restrict__ ctr, int s, int mode, double weight,
                   double alpha, double grid_r, double merge_limit, unsigned long long *__restrict__ stats)
{
    const DevState &st = *stp;
    const long long n = (long long)ctr->n[s];
    if ((long long)blockIdx.x * blockDim.x >= n) return;
    const long long cap = st.cap[s];
    __shared__ unsigned long long tally[4];   // copies, kept, deleted, dead slots
    if (threadIdx.x < 4) tally[threadIdx.x] = 0;
    __syncthreads();
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double *const *F = st.f[s];
    const double Z = i < n ? gp(F[0])[i] : kDead;
    if (!is_dead(Z)) {
        const double R = F[1][i];
        const double wq0 = F[7][i];
        double wq = wq0;
        const double ref = weight * ((1.0 - alpha) * R / grid_r + alpha);
        if (ref > 0.0) {
            const double Vz = F[2][i], Vr = F[3][i], Vt = F[4][i];
            if ((mode & 1) && wq / ref >= 2.0) {
                const double ratio = wq / ref;
                if (!(ratio - 1.0 <= (double)(cap - n))) {
                    atomicOr(&ctr->overflow, 1);            // more copies than free slots
                } else {
                    const unsigned long long m = (unsigned long long)(long long)ratio - 1ull;
                    const unsigned long long k0 = atomicAdd(&ctr->n_new[s], m);
                    const unsigned long long base = (unsigned long long)n + k0;
                    if (base + m > (unsigned long long)cap) {
                        // the reservation ran past the capacity: nothing is
                        // split, reserved slots inside it become dead slots
                        atomicOr(&ctr->overflow, 1);
                        unsigned long long nd = 0;
                        for (unsigned long long q = base; q < (unsigned long long)cap && q < base + m; ++q, ++nd)
                            F[0][q] = kDead;
                        if (nd) atomicAdd(&tally[3], nd);
                    } else {
                        const double Az = F[5][i], Ar = F[6][i];
                        const unsigned long long vtb = (unsigned long long)__double_as_longlong(Vt);
                        for (unsigned long long j = 1; j <= m; ++j) {
                            wq = wq - ref;
                            const unsigned long long q = base + j - 1;
                            F[0][q] = Z; F[1][q] = R; F[2][q] = Vz; F[3][q] = Vr;
                            F[4][q] = __longlong_as_double((long long)(vtb ^ j));
                            F[5][q] = Az; F[6][q] = Ar; F[7][q] = ref;
                        }
                        atomicAdd(&tally[0], m);
                    }
                }
            }
            if ((mode & 2) && wq < merge_limit * ref) {
                const StateKey sk = state_key(Z, R, Vz, Vr, Vt);
                Rng g = make_rng(sk, ctr->seed, ctr->step, kMergeTag | (uint32_t)s);
                if (g.next() < wq / ref) {
                    wq = ref;
                    atomicAdd(&tally[1], 1ull);
                } else {
                    F[0][i] = kDead;
                    atomicAdd(&tally[2], 1ull);
                    atomicAdd(&tally[3], 1ull);
                }
            }
            if (__double_as_longlong(wq) != __double_as_longlong(wq0)) F[7][i] = wq;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3 && tally[threadIdx.x]) atomicAdd(&stats[threadIdx.x], tally[threadIdx.x]);
    if (threadIdx.x == 3 && tally[3]) atomicAdd(&ctr->n_dead[s], tally[3]);
}

// Folds the copies appended by split_merge_kernel into n[s].
__global__ void rezone_fold_kernel(DevCounters *ctr, int s, long long cap)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const unsigned long long n = ctr->n[s] + ctr->n_new[s];
        ctr->n[s] = n > (unsigned long long)cap ? (unsigned long long)cap : n;
        ctr->n_new[s] = 0;
    }
}

// ------------------------------------------------------------ cell sort

// One block per tile of kSortTile slots.  The tile's cell keys are sorted in
// LDS first, so each run of equal cells in the tile costs one atomic on that
// cell's counter, and the run's particles go to consecutive slots of the
// cell's range.  Dead slots and slots past n take the sentinel key ncell and
// are dropped.  The order of particles within a cell depends on the order in
// which tiles reach the counters; nothing downstream depends on it (the
// collision draws are keyed by particle state, not by slot).
constexpr int kSortItems = 8;
constexpr int kSortTile = kBlock * kSortItems;          // default tile: 2048 slots, 256 threads
constexpr int kSortTileMax = 2 * kSortTile;             // 4096 slots, 512 threads

template <int SB> struct SortTypes {
    using Sort = hipcub::BlockRadixSort<unsigned, SB, kSortItems, unsigned short>;
    using Scan = hipcub::BlockScan<int, SB>;
};

template <int SB> struct SortTileLds {
    static constexpr int TILE = SB * kSortItems;
    union {
        typename SortTypes<SB>::Sort::TempStorage sort;
        typename SortTypes<SB>::Scan::TempStorage scan;
        unsigned short src[TILE];          // tile slot of each sorted position
    } tmp;
    unsigned key[TILE];                    // sorted keys, then destination slots
};

// SCATTER false: counts[cell] += particles of the cell (the histogram).
// SCATTER true: counts[cell] hands out ranks within the cell (zeroed before),
// and each particle's eight fields go to the species' other buffer set at
// offs[cell] + rank.  The grid covers the capacity; blocks past the
// population return at once, so a captured launch fits any population.
// SB threads sort a tile of SB * kSortItems slots: a larger tile gives
// longer runs of equal cells, so the scatter writes fewer partial lines.
template <bool SCATTER, int SB>
__global__ void __launch_bounds__(SB)
sort_tile_kernel(const DevState *__restrict__ stp, const DevCounters *__restrict__ ctr, int s,
                 unsigned long long *__restrict__ counts, const unsigned long long *__restrict__ offs)
{
    constexpr int TILE = SB * kSortItems;
    using TileSort = typename SortTypes<SB>::Sort;
    using TileScan = typename SortTypes<SB>::Scan;
    __shared__ SortTileLds<SB> L;
    // Scatter pass: one field of the tile at a time.  Before the fields, its
    // first TILE words hold the counting bins (or, after a radix sort, each
    // run's first rank in its cell) and the next TILE words the tile's first
    // rank in each cell of its span; the histogram pass needs the bins only.
    __shared__ double stage[SCATTER ? TILE : TILE / 2];
    unsigned *const base = reinterpret_cast<unsigned *>(stage);
    // The bounds of the tile's cells in z and r are reduced in the last four
    // words of L.key, which nothing writes before the counting (or the radix
    // sort's own barriers) has passed: separate variables would enlarge the
    // LDS block past what two 4096-slot (four 2048-slot) blocks share on a CU.
    unsigned &zlo_s = L.key[TILE - 4], &zhi_s = L.key[TILE - 3];
    unsigned &rlo_s = L.key[TILE - 2], &rhi_s = L.key[TILE - 1];
    const DevState &st = *stp;
    const long long n = (long long)ctr->n[s];
    const long long tbase = (long long)blockIdx.x * TILE;
    if (tbase >= n) return;
    const unsigned S = (unsigned)((st.nz - 1) * (st.nr - 1));
    const int t = threadIdx.x;
    const unsigned nzc = (unsigned)(st.nz - 1);     // cells per row (z fastest in the cell index)
    if (t == 0) { zlo_s = ~0u; zhi_s = 0u; rlo_s = ~0u; rhi_s = 0u; }
    // Z and R of the thread's slots, read together (slots past the
    // population read the tile's first slot and take the sentinel key); the
    // scatter pass moves these values as its first two fields, and loads the
    // third while the keys are counted.
    const int ntile = (int)min((long long)TILE, n - tbase);
    double zv[kSortItems], rv[kSortItems], nx[kSortItems];
    {
        const auto *fz = gp(st.f[s][0]) + tbase, *fr = gp(st.f[s][1]) + tbase;
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const int l = t + j * SB;
            zv[j] = __builtin_nontemporal_load(fz + (l < ntile ? l : 0));
            rv[j] = __builtin_nontemporal_load(fr + (l < ntile ? l : 0));
        }
        if (SCATTER) {
            const auto *f2 = gp(st.f[s][2]) + tbase;
#pragma unroll
            for (int j = 0; j < kSortItems; ++j) {
                const int l = t + j * SB;
                nx[j] = __builtin_nontemporal_load(f2 + (l < ntile ? l : 0));
            }
        }
    }
    // Cell coordinates (0-based) of the thread's slots; cr = ~0u marks a
    // slot without a live particle.
    unsigned cz[kSortItems], cr[kSortItems];
    unsigned short loc[kSortItems];
    unsigned zlo = ~0u, zhi = 0u, rlo = ~0u, rhi = 0u;
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int l = t + j * SB;
        cz[j] = 0u;
        cr[j] = ~0u;
        if (l < ntile && !is_dead(zv[j])) {
            cz[j] = (unsigned)(min(max((int)ceil(zv[j]), 1), st.nz - 1) - 1);
            cr[j] = (unsigned)(min(max((int)ceil(rv[j]), 1), st.nr - 1) - 1);
            zlo = min(zlo, cz[j]); zhi = max(zhi, cz[j]);
            rlo = min(rlo, cr[j]); rhi = max(rhi, cr[j]);
        }
        loc[j] = (unsigned short)l;
    }
    const int lane = t & 63;
    for (int o = 32; o > 0; o >>= 1) {
        zlo = min(zlo, (unsigned)__shfl_xor((int)zlo, o)); zhi = max(zhi, (unsigned)__shfl_xor((int)zhi, o));
        rlo = min(rlo, (unsigned)__shfl_xor((int)rlo, o)); rhi = max(rhi, (unsigned)__shfl_xor((int)rhi, o));
    }
    __syncthreads();                        // bounds initialised
    if (lane == 0) {
        atomicMin(&zlo_s, zlo); atomicMax(&zhi_s, zhi);
        atomicMin(&rlo_s, rlo); atomicMax(&rhi_s, rhi);
    }
    __syncthreads();
    zlo = zlo_s; zhi = zhi_s; rlo = rlo_s; rhi = rhi_s;
    if (rlo > rhi) return;                  // no live particle in the tile (uniform)
    // The tile's cells lie in the box [zlo, zhi] x [rlo, rhi]: between sorts
    // a few dozen cells in z by a few rows in r, a few hundred cells in all
    // though their cell indices span several rows of the grid.  Numbered
    // row by row within the box, they keep the order of the cell index.
    const unsigned W = zhi - zlo + 1u, H = rhi - rlo + 1u;
    unsigned key[kSortItems];
    if ((unsigned long long)W * H <= (unsigned long long)TILE) {
        const unsigned span = W * H;        // box index of the sentinel
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) key[j] = cr[j] != ~0u ? (cr[j] - rlo) * W + (cz[j] - zlo) : span;
        // Counted in LDS, one bin per cell of the box: a wave's 64 slots
        // (consecutive) fall into a few runs of equal cells, and each run
        // takes its ranks with one LDS atomic; each nonzero bin then costs
        // one atomic on its cell's counter.  The bins' prefix sums place
        // the runs in the tile's sorted order.  (Tiles whose box holds
        // more than TILE cells -- the products appended since the last
        // sort -- are radix-sorted below.)
        unsigned *bins = base;
        for (int k = t; k < (int)span; k += SB) bins[k] = 0u;
        __syncthreads();
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        unsigned rank[kSortItems];
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const unsigned rel = key[j];
            const unsigned kp = __shfl_up(rel, 1);
            const unsigned long long starts = __ballot(lane == 0 || kp != rel);
            const int first = 63 - __clzll((long long)(starts & upto));
            const unsigned long long above = starts & ~upto;
            const int last = (above ? __ffsll((long long)above) - 1 : 64) - 1;
            unsigned old = 0u;
            if (lane == last && rel < span) old = atomicAdd(&bins[rel], (unsigned)(lane - first + 1));
            old = __shfl(old, last);
            rank[j] = old + (unsigned)(lane - first);
        }
        __syncthreads();                    // every bin counted
        unsigned c8[kSortItems];
        unsigned tsum = 0u;
#pragma unroll
        for (int q = 0; q < kSortItems; ++q) {
            const int k = t * kSortItems + q;
            c8[q] = k < (int)span ? bins[k] : 0u;
            tsum += c8[q];
        }
        // the destination of each bin's first particle: its cell's offset
        // plus the tile's first rank within the cell (scatter pass: kept in
        // the staging buffer, which the fields use only later)
        unsigned *gb = SCATTER ? base + TILE : nullptr;
        unsigned bz = (unsigned)(t * kSortItems) % W, br = (unsigned)(t * kSortItems) / W;
#pragma unroll
        for (int q = 0; q < kSortItems; ++q) {
            if (c8[q]) {
                const int k = t * kSortItems + q;
                const unsigned cell = (zlo + bz) + (rlo + br) * nzc;
                const unsigned long long g = __hip_atomic_fetch_add(gp(&counts[cell]), (unsigned long long)c8[q],
                                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (SCATTER) gb[k] = (unsigned)(gp(offs)[cell] + g);
            }
            if (++bz == W) { bz = 0u; ++br; }
        }
        if (!SCATTER) return;
        int excl = 0, total = 0;
        TileScan(L.tmp.scan).ExclusiveSum((int)tsum, excl, total);
        __syncthreads();                    // scan storage free; bins read
        unsigned run = (unsigned)excl;
#pragma unroll
        for (int q = 0; q < kSortItems; ++q) {
            const int k = t * kSortItems + q;
            if (k < (int)span) bins[k] = run;   // first sorted position of the bin
            run += c8[q];
        }
        for (int p = total + t; p < TILE; p += SB) {   // positions past the tile's particles
            L.key[p] = 0xffffffffu;
            L.tmp.src[p] = 0;
        }
        __syncthreads();                    // starts written
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            if (key[j] < span) {
                const unsigned p = bins[key[j]] + rank[j];
                L.tmp.src[p] = (unsigned short)(t + j * SB);
                L.key[p] = gb[key[j]] + rank[j];
            }
        }
    } else {
        // radix-sorted on the cell index relative to the box's first cell,
        // on only the bits up to the box's last cell (the sentinel, one past
        // it, takes the last place)
        const unsigned kmin = zlo + rlo * nzc;
        const unsigned span = zhi + rhi * nzc - kmin + 1u;
        int nbits = 0;
        while (nbits < 32 && (span >> nbits) != 0) ++nbits;
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) key[j] = cr[j] != ~0u ? cz[j] + cr[j] * nzc - kmin : span;
        TileSort(L.tmp.sort).Sort(key, loc, 0, nbits);        // blocked: thread t holds 8t .. 8t+7
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) key[j] = key[j] < span ? key[j] + kmin : S;
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) L.key[t * kSortItems + j] = key[j];
        __syncthreads();
        int head[kSortItems], start[kSortItems];
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const int p = t * kSortItems + j;
            head[j] = (p == 0 || L.key[p - 1] != key[j]) ? p : 0;
        }
        TileScan(L.tmp.scan).InclusiveScan(head, start, hipcub::Max());   // first position of each run
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const int p = t * kSortItems + j;
            const bool run_end = p == TILE - 1 || L.key[p + 1] != key[j];
            if (run_end && key[j] < S) {
                const unsigned long long len = (unsigned long long)(p - start[j] + 1);
                const unsigned long long b =
                    __hip_atomic_fetch_add(gp(&counts[key[j]]), len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (SCATTER) base[start[j]] = (unsigned)b;
            }
        }
        if (!SCATTER) return;
        __syncthreads();                        // base[] written; tmp.scan and key[] reads done
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const int p = t * kSortItems + j;
            L.key[p] = key[j] < S ? (unsigned)(gp(offs)[key[j]] + base[start[j]] + (unsigned)(p - start[j])) : 0xffffffffu;
            L.tmp.src[p] = loc[j];
        }
    }
    // Field by field: the tile's values are read coalesced (non-temporal:
    // read once) into LDS, then written in sorted order, neighbouring lanes
    // to neighbouring slots of a run.  Reading them straight from HBM in
    // sorted order costs a cache line per lane.  Each thread keeps the
    // destinations and source slots of its sorted positions in registers,
    // and loads the next field while the current one is written out (Z and
    // R are the key phase's values, the third field was loaded then;
    // sorted positions without a particle write to the spare slot past the
    // capacity, so no load or store is conditional and the waits for the
    // loads need not wait for the stores).  The barriers only order LDS: no
    // memory operation is waited for there.
    __syncthreads();                        // key/src written
    unsigned dd[kSortItems];
    int ss[kSortItems];
    const unsigned spare = (unsigned)st.cap[s];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const unsigned d = L.key[t + j * SB];
        dd[j] = d != 0xffffffffu ? d : spare;
        ss[j] = L.tmp.src[t + j * SB];
    }
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        __builtin_amdgcn_s_waitcnt(0xC07F);       // lgkmcnt(0): this thread's LDS reads of stage are done
        __builtin_amdgcn_s_barrier();            // ... and every other thread's: stage may be overwritten
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) stage[t + j * SB] = f == 0 ? zv[j] : f == 1 ? rv[j] : nx[j];
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_s_barrier();            // stage written
        if (f >= 2 && f < 7) {
            const auto *sf = gp(st.f[s][f + 1]) + tbase;
#pragma unroll
            for (int j = 0; j < kSortItems; ++j) {
                const int l = t + j * SB;
                nx[j] = __builtin_nontemporal_load(sf + (l < ntile ? l : 0));
            }
        }
        double v[kSortItems];
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) v[j] = stage[ss[j]];
        auto *df = gp(st.alt[s][f]);
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) df[dd[j]] = v[j];
    }
}

// After the scatter: the sorted set becomes species s's current buffer set.
__global__ void swap_buffers_kernel(DevState *st, int s)
{
    if (threadIdx.x < 8) {
        const int f = threadIdx.x;
        double *a = st->f[s][f];
        st->f[s][f] = st->alt[s][f];
        st->alt[s][f] = a;
    }
}

__global__ void clear_u64_kernel(unsigned long long *a, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0ull;
}

// Exclusive scan of n counts: per-block scan + block totals, then offsets.
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock)
scan_block_kernel(const unsigned long long *__restrict__ in, unsigned long long *__restrict__ out,
                  unsigned long long *__restrict__ partial, size_t n)
{
    __shared__ unsigned long long wsum[kScanBlock / 64];
    const size_t i = (size_t)blockIdx.x * kScanBlock + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long v = (i < n) ? in[i] : 0ull;
    unsigned long long x = v;
    for (int o = 1; o < 64; o <<= 1) {
        unsigned long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        unsigned long long a = (lane < kScanBlock / 64) ? wsum[lane] : 0ull;
        for (int o = 1; o < 64; o <<= 1) {
            unsigned long long y = __shfl_up(a, o);
            if (lane >= o) a += y;
        }
        if (lane < kScanBlock / 64) wsum[lane] = a;
    }
    __syncthreads();
    const unsigned long long before = (w > 0) ? wsum[w - 1] : 0ull;
    if (i < n) out[i] = before + x - v;
    if (threadIdx.x == kScanBlock - 1) partial[blockIdx.x] = before + x;
}

// add[b] = exclusive prefix of the block totals
__global__ void scan_add_kernel(unsigned long long *__restrict__ out, const unsigned long long *__restrict__ add,
                                size_t n)
{
    const size_t i = (size_t)blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) out[i] += add[blockIdx.x];
}

// Deposit-window anchors of a species' tiles in the order a sort left: each
// tile's window is placed by its first particle, as the particle pass does
// for the pass after it.  anc points at the species' first tile.
__global__ void anchor_kernel(const DevState *__restrict__ stp, const DevCounters *__restrict__ ctr, int s,
                              int *__restrict__ anc, long long max_tiles, int arows)
{
    const long long n = (long long)ctr->n[s];
    const long long nt = min((n + kTile - 1) / kTile, max_tiles);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const double *__restrict__ Z = stp->f[s][0];
    const double *__restrict__ R = stp->f[s][1];
    const double z = Z[t * kTile], r = R[t * kTile];
    if (is_dead(z) || !(z > 0.0) || !(r >= 0.0) || z >= 1e9 || r >= 1e9) return;
    anc[2 * t] = (int)ceil(z) - kAnchorZ;
    anc[2 * t + 1] = (int)ceil(r) - arows;
}

__global__ void set_count_kernel(DevCounters *ctr, int s, const unsigned long long *offs,
                                 const unsigned long long *hist, size_t ncell)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctr->n[s] = offs[ncell - 1] + hist[ncell - 1];
        ctr->n_dead[s] = 0;
    }
}

}  // namespace

// ------------------------------------------------- no-allocation scratch
// Accumulators of ipm_moments_noalloc: five node grids of at most
// kScratchNodes nodes (40 MiB).  Each GPU the library runs on holds its own
// copy for the life of the process; every call on a GPU shares it, so calls
// that overlap on one GPU race.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchNodes = (size_t)1 << 20;
static __device__ __align__(256) double scratch[5 * kScratchNodes];

// ====================================================================
//                              host side
// ====================================================================
#include "ipm_ctx.hpp"

// Device address of the scratch array on device dev (the current device),
// looked up once per device and kept.
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64] = {};
static int scratch_address(int dev, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= 64) return fail(IPM_EINVAL, "device %d beyond the scratch table", dev);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(IPM_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = (double *)g_scratch_ptr[dev];
    return IPM_OK;
}

// Per-table constants of Updater and VelMaxInit (R/src/mc/TypeMCC.f90:60-92,
// R/src/pic/particle/ParticleOne.f90:76-95) that depend only on the species
// and gas masses and the gas temperature, formed by the same expressions the
// device would evaluate per collision.
static void table_consts(ipm_ctx *c)
{
    for (int s = 0; s < c->ns; ++s)
        for (int g = 0; g < c->ng; ++g) {
            DevTable &t = c->h.tab[s * c->ng + g];
            if (t.model == 0) continue;
            const double sm = c->h.mass[s], gm = t.gas_mass;
            t.miu = sm * gm / (sm + gm);
            t.mass_ratio = gm / (sm + gm);
            t.beta = 1.0 / (kKB * t.gas_temperature);
            // VelMaxInit's speed sqrt(-3 log(R) / Beta / Mass) with R >= 2^-53
            // (-log R <= 36.737), with margin
            t.nb_vmax = sqrt(3.0 * 36.75 * kKB * t.gas_temperature / gm) * (1.0 + 1e-9);
            t.nb_ecoef = 0.5 * t.miu / kJTOEV;
        }
}

// The null-event screen's bound table of a model 1 or 2 table (layout in
// DevTable): per energy bin the largest probability any reaction can
// interpolate to there, raised by 1e-12 for rounding, then the maxima over
// runs of 2^L bins.  Tables with a value that is not finite get none.
static std::vector<double> null_bound_table(int nreaction, int nsigma, const double *prob)
{
    const int Ns = nsigma, Nr = nreaction;
    std::vector<double> rowmax(Ns);
    for (int i = 0; i < Ns; ++i) {
        double m = 0.0;
        for (int j = 0; j < Nr; ++j) {
            const double v = prob[(size_t)i * Nr + j];
            if (!std::isfinite(v)) return {};
            m = std::max(m, v);
        }
        rowmax[i] = m;
    }
    int levels = 1;
    while ((2 << (levels - 1)) <= Ns + 1) ++levels;
    std::vector<double> tab((size_t)levels * (Ns + 1), 2.0);
    for (int b = 0; b <= Ns; ++b) {
        const double m = b == 0 ? rowmax[0] : b == Ns ? rowmax[Ns - 1] : std::max(rowmax[b - 1], rowmax[b]);
        tab[b] = m * (1.0 + 1e-12);
    }
    for (int L = 1; L < levels; ++L) {
        const int h = 1 << (L - 1);
        for (int b = 0; b + 2 * h - 1 <= Ns; ++b)
            tab[(size_t)L * (Ns + 1) + b] =
                std::max(tab[(size_t)(L - 1) * (Ns + 1) + b], tab[(size_t)(L - 1) * (Ns + 1) + b + h]);
    }
    return tab;
}

// Collision-queue regions: species s gets ceil(cap_s/1024) tile regions of
// R_s = min(1024, ceil(1024 * min(1, 2 * sum_g CollisionRatio)) + 64)
// entries, at least 4 standard deviations above the binomial mean of a
// tile's selections for any ratio (a larger count stops the step with
// IPM_ECAPACITY).  Species without MCC tables get none.
static int ensure_queue(ipm_ctx *c)
{
    long long R[IPM_NS_MAX] = {}, nt[IPM_NS_MAX] = {};
    long long need = 0, ntiles = 0;
    for (int s = 0; s < c->ns; ++s) {
        double r = 0.0;
        bool any = false;
        for (int g = 0; g < c->ng; ++g)
            if (c->h.tab[s * c->ng + g].model != 0) { r += c->h.tab[s * c->ng + g].ratio; any = true; }
        nt[s] = c->species_set[s] ? (c->h.cap[s] + kTile - 1) / kTile : 0;
        if (any) {
            const double f = std::min(1.0, 2.0 * r);
            R[s] = std::min<long long>(kTile, (long long)ceil((double)kTile * f) + 64);
        }
        c->h.q_weight[s] = any ? (double)kTile * std::min(1.0, r) + 1e-9 : 0.0;
        c->h.q_tile_cap[s] = R[s];
        c->h.q_region[s] = need;
        c->h.q_cbase[s] = ntiles;
        need += nt[s] * R[s];
        ntiles += nt[s];
    }
    need = std::max<long long>(need, 1);
    ntiles = std::max<long long>(ntiles, 1);
    if (need > c->q_cap || ntiles > c->q_ntiles) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->qbuf) HIPCHK(hipFree(c->qbuf));
        if (c->q_tile_count) HIPCHK(hipFree(c->q_tile_count));
        if (c->tanc) HIPCHK(hipFree(c->tanc));
        c->tanc = nullptr;
        c->qbuf = nullptr; c->q_tile_count = nullptr; c->q_cap = 0; c->q_ntiles = 0;
        HIPCHK(hipMalloc(&c->qbuf, sizeof(double) * kRecD * (size_t)need));
        HIPCHK(hipMalloc(&c->q_tile_count, sizeof(unsigned) * (size_t)ntiles));
        HIPCHK(hipMalloc(&c->tanc, sizeof(int) * 2 * (size_t)ntiles));
        HIPCHK(hipMemset(c->tanc, 0, sizeof(int) * 2 * (size_t)ntiles));
        HIPCHK(hipMemset(c->q_tile_count, 0, sizeof(unsigned) * (size_t)ntiles));
        c->q_cap = need;
        c->q_ntiles = ntiles;
        drop_graphs(c);
    }
    if (!c->p_stage) {
        const size_t G = (size_t)c->num_cu * kCollideBlocksPerCuMax;
        HIPCHK(hipMalloc(&c->p_stage, sizeof(double) * 8 * kProdStage * G));
        HIPCHK(hipMalloc(&c->p_meta, sizeof(unsigned) * kProdStage * G));
        HIPCHK(hipMalloc(&c->p_cnt, sizeof(unsigned) * IPM_NS_MAX * G));
        HIPCHK(hipMemset(c->p_cnt, 0, sizeof(unsigned) * IPM_NS_MAX * G));
    }
    c->h.c_grid = c->num_cu * c->collide_bpc;
    if (c->h.c_grid < c->ns)
        return fail(IPM_EINVAL, "collision grid of %d blocks is smaller than the %d species", c->h.c_grid, c->ns);
    c->h.q_rec = c->qbuf;
    c->h.tile_anchor = c->tanc;
    c->h.p_stage = c->p_stage;
    c->h.p_meta = c->p_meta;
    c->h.p_cnt = c->p_cnt;
    c->h.q_tile_count = c->q_tile_count;
    return IPM_OK;
}

// blocks_per_cu > 0: persistent grid of that many blocks per CU looping over
// tiles.  blocks_per_cu == 0: flat grid of one block per tile, sized from the
// species capacities so that a captured graph covers any population; blocks
// past the live tiles return at once.
static int grid_blocks(ipm_ctx *c)
{
    if (c->blocks_per_cu > 0) {
        int g = c->num_cu * c->blocks_per_cu;
        return g < 8 ? 8 : g;
    }
    long long tmax = 0;
    for (int s = 0; s < c->ns; ++s) tmax += (c->h.cap[s] + kTile - 1) / kTile;
    tmax = ((tmax + 7) / 8) * 8;
    return (int)std::max<long long>(8, std::min<long long>(tmax, 0x7fffffffLL));
}

template <int MODE, int BLOCK, int WINR>
static void launch_shape(ipm_ctx *c, int sel)
{
    hipLaunchKernelGGL((particle_kernel<MODE, BLOCK, WINR>), dim3(grid_blocks(c)), dim3(BLOCK), 0, c->stream,
                       c->d_state, c->d_ctr, sel);
}

// The fused pass (every mode bit set) runs the tuned shape; the separate
// entry points run the default one.
template <int MODE>
static int launch_particles(ipm_ctx *c, int sel)
{
    if constexpr ((MODE & ~M_REZONE) == (M_PUSH | M_ABSORB | M_DEPOSIT | M_MCC)) {
        const int b = c->pk_block, w = c->pk_winr;
        if (b == 128 && w == 24) launch_shape<MODE, 128, 24>(c, sel);
        else if (b == 512 && w == 24) launch_shape<MODE, 512, 24>(c, sel);
        else if (b == 256 && w == 16) launch_shape<MODE, 256, 16>(c, sel);
        else if (b == 256 && w == 32) launch_shape<MODE, 256, 32>(c, sel);
        else if (b == 256 && w == 24) launch_shape<MODE, 256, 24>(c, sel);
        else return fail(IPM_EINVAL, "no particle-pass variant for %d threads and a %d-row window", b, w);
    } else {
        launch_shape<MODE, kBlock, kWinR>(c, sel);
    }
    HIPCHK(hipGetLastError());
    return IPM_OK;
}

// fold: also fold the products into the counts (else the caller's
// grid_kernel launch does it)
static int launch_collide(ipm_ctx *c, bool fold)
{
    const unsigned G = (unsigned)c->h.c_grid;
    hipLaunchKernelGGL(collide_kernel, dim3(G), dim3(kBlock), 0, c->stream, c->d_state, c->d_ctr);
    hipLaunchKernelGGL(collide_append_kernel, dim3(G), dim3(kBlock), 0, c->stream, c->d_state, c->d_ctr);
    if (fold)
        hipLaunchKernelGGL(collide_fold_kernel, dim3(1), dim3(kBlock), 0, c->stream, c->d_state, c->d_ctr);
    HIPCHK(hipGetLastError());
    return IPM_OK;
}

static int launch_grid(ipm_ctx *c, int s0, int s1, int sum, int fold = 0)
{
    const size_t nn = (size_t)c->nz * c->nr;
    const size_t ngh = (size_t)(c->nz + 2) * (c->nr + 2);
    const size_t nt = std::max(nn, ngh);
    hipLaunchKernelGGL(grid_kernel, dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                       c->d_state, c->d_ctr, c->inv_vol, c->d_chifac, c->rho_one, c->chi_one, c->rho_tot, c->chi_tot,
                       s0, s1, sum, fold);
    HIPCHK(hipGetLastError());
    return IPM_OK;
}

// Enqueues the cell sort of species s0..s1-1 (device work only, no host
// synchronization, capturable): per species, histogram, scan, scatter into
// the other buffer set, new count, buffer swap on the device, window anchors.
// Grids are sized from the capacities.  The caller mirrors the swap on the
// host with sort_mirror once the work is enqueued (or replayed).
static int sort_enqueue(ipm_ctx *c, int s0, int s1)
{
    const size_t ncell = (size_t)(c->nz - 1) * (c->nr - 1);
    const unsigned nblk = (unsigned)((ncell + kScanBlock - 1) / kScanBlock);
    if (nblk > kScanBlock) return fail(IPM_EINVAL, "grid too large for the two-level scan");
    const unsigned cblk = (unsigned)((ncell + kBlock - 1) / kBlock);
    for (int sp = s0; sp < s1; ++sp) {
        if (!c->species_set[sp]) continue;
        hipLaunchKernelGGL(clear_u64_kernel, dim3(cblk), dim3(kBlock), 0, c->stream, c->hist, ncell);
        hipLaunchKernelGGL(clear_u64_kernel, dim3(cblk), dim3(kBlock), 0, c->stream, c->cursor, ncell);
        const bool wide = c->sort_block == 2 * kBlock;
        const size_t tile = wide ? kSortTileMax : kSortTile;
        const unsigned pblk = (unsigned)(((size_t)c->h.cap[sp] + tile - 1) / tile);
        const dim3 sblk(wide ? 2 * kBlock : kBlock);
        const unsigned long long *no_offs = nullptr;
        if (wide)
            hipLaunchKernelGGL((sort_tile_kernel<false, 2 * kBlock>), dim3(pblk), sblk, 0, c->stream, c->d_state,
                               c->d_ctr, sp, c->hist, no_offs);
        else
            hipLaunchKernelGGL((sort_tile_kernel<false, kBlock>), dim3(pblk), sblk, 0, c->stream, c->d_state, c->d_ctr,
                               sp, c->hist, no_offs);
        hipLaunchKernelGGL(scan_block_kernel, dim3(nblk), dim3(kScanBlock), 0, c->stream, c->hist, c->offs, c->partial, ncell);
        hipLaunchKernelGGL(scan_block_kernel, dim3(1), dim3(kScanBlock), 0, c->stream, c->partial, c->partial2,
                           c->partial + nblk, (size_t)nblk);
        hipLaunchKernelGGL(scan_add_kernel, dim3(nblk), dim3(kScanBlock), 0, c->stream, c->offs, c->partial2, ncell);
        if (wide)
            hipLaunchKernelGGL((sort_tile_kernel<true, 2 * kBlock>), dim3(pblk), sblk, 0, c->stream, c->d_state,
                               c->d_ctr, sp, c->cursor, c->offs);
        else
            hipLaunchKernelGGL((sort_tile_kernel<true, kBlock>), dim3(pblk), sblk, 0, c->stream, c->d_state, c->d_ctr,
                               sp, c->cursor, c->offs);
        hipLaunchKernelGGL(set_count_kernel, dim3(1), dim3(64), 0, c->stream, c->d_ctr, sp, c->offs, c->hist, ncell);
        hipLaunchKernelGGL(swap_buffers_kernel, dim3(1), dim3(64), 0, c->stream, c->d_state, sp);
        if (c->tanc) {
            const long long max_tiles = ((long long)c->h.cap[sp] + kTile - 1) / kTile;
            if (c->h.q_cbase[sp] + max_tiles <= c->q_ntiles) {
                const unsigned ablk = (unsigned)((max_tiles + kBlock - 1) / kBlock);
                hipLaunchKernelGGL(anchor_kernel, dim3(ablk), dim3(kBlock), 0, c->stream, c->d_state, c->d_ctr, sp,
                                   c->tanc + 2 * c->h.q_cbase[sp], max_tiles, anchor_rows(c->pk_winr));
            }
        }
        HIPCHK(hipGetLastError());
    }
    return IPM_OK;
}

static void sort_mirror(ipm_ctx *c, int s0, int s1)
{
    for (int sp = s0; sp < s1; ++sp)
        if (c->species_set[sp]) set_buffers(c, sp, 1 - c->cur[sp]);
}

// The sort of every species as a captured graph (captured on first use).
static int sort_graph_launch(ipm_ctx *c)
{
    if (!c->sort_graph) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = sort_enqueue(c, 0, c->ns);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return fail(IPM_EHIP, "graph capture failed: %s", hipGetErrorString(e));
        HIPCHK(hipGraphInstantiate(&c->sort_graph, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
    }
    HIPCHK(hipGraphLaunch(c->sort_graph, c->stream));
    sort_mirror(c, 0, c->ns);
    return IPM_OK;
}

extern "C" {

const char *ipm_last_error(void) { return g_err; }

int ipm_create(ipm_ctx **out, int nz, int nr, double dx, double dt, int ns, int ng)
{
    if (!out) return fail(IPM_EINVAL, "null output pointer");
    *out = nullptr;
    if (nz < 3 || nr < 3 || !(dx > 0) || !(dt > 0) || ns < 1 || ns > IPM_NS_MAX || ng < 0 ||
        ng > IPM_NG_MAX)
        return fail(IPM_EINVAL, "invalid grid/species: nz=%d nr=%d dx=%g dt=%g ns=%d ng=%d", nz, nr, dx,
                    dt, ns, ng);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(IPM_ENODEV, "no HIP device");
    ipm_ctx *c = new ipm_ctx();
    HIPCHK(hipGetDevice(&c->device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, c->device));
    c->num_cu = prop.multiProcessorCount;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        const int rc = scratch_address(c->device, &c->scratch);
        if (rc) { ipm_destroy(c); return rc; }
    }
    c->nz = nz; c->nr = nr; c->ns = ns; c->ng = ng; c->dx = dx; c->dt = dt;
    DevState &h = c->h;
    h.nz = nz; h.nr = nr; h.ns = ns; h.ng = ng; h.dx = dx; h.dt = dt;
    h.vfactor = dx / dt;               // ParticleBundle.f90:346
    h.inv_vfactor = 1.0 / h.vfactor;   // MCCPublic.F90:90 multiplies by 1/VFactor
    h.null_screen = 1;
    const size_t nn = (size_t)nz * nr;
    HIPCHK(hipMalloc(&c->d_state, sizeof(DevState)));
    HIPCHK(hipMalloc(&c->d_ctr, sizeof(DevCounters)));
    HIPCHK(hipMemset(c->d_ctr, 0, sizeof(DevCounters)));
    HIPCHK(hipMalloc(&c->raw_all, sizeof(double) * nn * ns));
    HIPCHK(hipMemset(c->raw_all, 0, sizeof(double) * nn * ns));
    HIPCHK(hipMalloc(&c->rho_one, sizeof(double) * nn * ns));
    HIPCHK(hipMemset(c->rho_one, 0, sizeof(double) * nn * ns));
    HIPCHK(hipMalloc(&c->chi_one, sizeof(double) * nn * ns));
    HIPCHK(hipMemset(c->chi_one, 0, sizeof(double) * nn * ns));
    HIPCHK(hipMalloc(&c->rho_tot, sizeof(double) * nn));
    HIPCHK(hipMalloc(&c->chi_tot, sizeof(double) * (nz + 2) * (nr + 2)));
    HIPCHK(hipMalloc(&c->d_chifac, sizeof(double) * IPM_NS_MAX));
    HIPCHK(hipMalloc(&c->Ez, sizeof(double) * (nz + 1) * nr));
    HIPCHK(hipMalloc(&c->Er, sizeof(double) * nz * (nr + 1)));
    HIPCHK(hipMemset(c->Ez, 0, sizeof(double) * (nz + 1) * nr));
    HIPCHK(hipMemset(c->Er, 0, sizeof(double) * nz * (nr + 1)));
    HIPCHK(hipMalloc(&c->sigma, sizeof(double) * nn));
    HIPCHK(hipMemset(c->sigma, 0, sizeof(double) * nn));
    HIPCHK(hipMalloc(&c->mcc_counts, sizeof(unsigned long long) * IPM_NS_MAX * IPM_NG_MAX * 128));
    HIPCHK(hipMemset(c->mcc_counts, 0, sizeof(unsigned long long) * IPM_NS_MAX * IPM_NG_MAX * 128));
    const size_t ncell = (size_t)(nz - 1) * (nr - 1);
    const size_t nblk = (ncell + kScanBlock - 1) / kScanBlock;
    HIPCHK(hipMalloc(&c->hist, sizeof(unsigned long long) * ncell));
    HIPCHK(hipMalloc(&c->offs, sizeof(unsigned long long) * ncell));
    HIPCHK(hipMalloc(&c->cursor, sizeof(unsigned long long) * ncell));
    HIPCHK(hipMalloc(&c->partial, sizeof(unsigned long long) * (nblk + 1)));
    HIPCHK(hipMalloc(&c->partial2, sizeof(unsigned long long) * (nblk + 1)));
    for (int s = 0; s < ns; ++s) h.raw[s] = c->raw_all + (size_t)s * nn;
    h.Ez = c->Ez; h.Er = c->Er; h.sigma = c->sigma; h.mcc_counts = c->mcc_counts;
    // no geometry yet: a 1-metal, all-vacuum placeholder keeps pointers valid
    h.ndiel = 0; h.nmetal = 0;
    int rc = sync_state(c);
    if (rc) { ipm_destroy(c); return rc; }
    *out = c;
    return IPM_OK;
}

int ipm_destroy(ipm_ctx *c)
{
    if (!c) return IPM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)window_release(c);
    (void)recomb_release(c);
    drop_graphs(c);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; ++b)
        for (int s = 0; s < IPM_NS_MAX; ++s)
            if (c->buf[b][s]) (void)hipFree(c->buf[b][s]);
    for (auto &v : c->tab_mem)
        for (void *p : v) (void)hipFree(p);
    void *ptrs[] = {c->d_state, c->d_ctr, c->raw_all, c->rho_one, c->chi_one, c->rho_tot, c->chi_tot,
                    c->d_chifac, c->inv_vol, c->cell_material, c->hedge, c->vedge, c->mat_type, c->row_vac,
                    c->metal_charge, c->sigma, c->Ez, c->Er, c->mcc_counts, c->hist, c->offs,
                    c->cursor, c->partial, c->partial2, c->qbuf, c->q_tile_count, c->tanc, c->p_stage,
                    c->p_meta, c->p_cnt,
                    c->mom, c->phi, c->eps_cell, c->ptype, c->sol, c->rezone};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return IPM_OK;
}

int ipm_set_species(ipm_ctx *c, int s, double charge, double mass, int64_t capacity)
{
    if (!c || s < 0 || s >= c->ns) return fail(IPM_EINVAL, "invalid species index %d", s);
    if (!(mass > 0) || capacity < 1) return fail(IPM_EINVAL, "invalid mass/capacity for species %d", s);
    if (capacity > IPM_CAPACITY_MAX)
        return fail(IPM_EINVAL, "capacity %lld of species %d exceeds IPM_CAPACITY_MAX (%lld): the particle pass "
                    "addresses each field with 32-bit byte offsets", (long long)capacity, s, (long long)IPM_CAPACITY_MAX);
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int b = 0; b < 2; ++b) {
        if (c->buf[b][s]) HIPCHK(hipFree(c->buf[b][s]));
        c->buf[b][s] = nullptr;
    }
    const size_t cap = field_stride(capacity);
    for (int b = 0; b < 2; ++b) HIPCHK(hipMalloc(&c->buf[b][s], sizeof(double) * 8 * cap));
    c->h.cap[s] = (long long)capacity;
    set_buffers(c, s, 0);
    drop_graphs(c);
    c->h.charge[s] = charge;
    c->h.mass[s] = mass;
    // MoveESIm: EFactor = Charge/Mass*dt/(dx/dt)*0.5 (ParticleBundle.f90:682)
    c->h.efactor[s] = charge / mass * c->dt / (c->dx / c->dt) * 0.5;
    // ResetFieldOne: ChiFactor = 0.5*Charge/Mass*dt*dt/Epsilon (FieldOne.f90:53)
    c->chifac[s] = 0.5 * charge / mass * c->dt * c->dt / kEPS0;
    HIPCHK(hipMemcpy(c->d_chifac, c->chifac, sizeof(double) * IPM_NS_MAX, hipMemcpyHostToDevice));
    c->species_set[s] = true;
    table_consts(c);
    DevCounters ctr;
    HIPCHK(hipMemcpy(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost));
    ctr.n[s] = 0; ctr.n_new[s] = 0; ctr.n_dead[s] = 0;
    HIPCHK(hipMemcpy(c->d_ctr, &ctr, sizeof ctr, hipMemcpyHostToDevice));
    int rc = ensure_queue(c);
    if (rc) return rc;
    return sync_state(c);
}

int ipm_set_geometry(ipm_ctx *c, const double *inv_volume, const int *cell_material, const int *hedge_type,
                     const int *vedge_type, int ndiel, int nmetal, const int *material_type)
{
    if (!c || !inv_volume || !cell_material || !hedge_type || !vedge_type || !material_type || ndiel < 0 ||
        nmetal < 0)
        return fail(IPM_EINVAL, "invalid geometry arguments");
    const int nz = c->nz, nr = c->nr;
    const size_t ncell = (size_t)(nz - 1) * (nr - 1);
    for (size_t k = 0; k < ncell; ++k)
        if (cell_material[k] < -ndiel || cell_material[k] > nmetal)
            return fail(IPM_EINVAL, "cell_material[%zu]=%d outside [-%d, %d]", k, cell_material[k], ndiel, nmetal);
    HIPCHK(hipStreamSynchronize(c->stream));
    void *old[] = {c->inv_vol, c->cell_material, c->hedge, c->vedge, c->mat_type, c->row_vac, c->metal_charge};
    for (void *p : old)
        if (p) HIPCHK(hipFree(p));
    const size_t nn = (size_t)nz * nr;
    HIPCHK(hipMalloc(&c->inv_vol, sizeof(double) * nn));
    HIPCHK(hipMemcpy(c->inv_vol, inv_volume, sizeof(double) * nn, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->cell_material, sizeof(int) * ncell));
    HIPCHK(hipMemcpy(c->cell_material, cell_material, sizeof(int) * ncell, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->hedge, sizeof(int) * nz * (nr - 1)));
    HIPCHK(hipMemcpy(c->hedge, hedge_type, sizeof(int) * nz * (nr - 1), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->vedge, sizeof(int) * (nz - 1) * nr));
    HIPCHK(hipMemcpy(c->vedge, vedge_type, sizeof(int) * (nz - 1) * nr, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->mat_type, sizeof(int) * (ndiel + nmetal + 1)));
    HIPCHK(hipMemcpy(c->mat_type, material_type, sizeof(int) * (ndiel + nmetal + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->metal_charge, sizeof(double) * (nmetal + 1)));
    HIPCHK(hipMemset(c->metal_charge, 0, sizeof(double) * (nmetal + 1)));
    // Per cell row, the longest run of cells that are neither metal nor
    // dielectric (absorb_one's fast path); the entries of rows nr - 1 .. nr + 1
    // stay empty (the particle pass reads four rows from any row below nr - 1).
    std::vector<int2> rv((size_t)nr + 2, int2{0, 0});
    for (int j = 0; j < nr - 1; ++j) {
        int best0 = 0, best1 = 0, run0 = 0;
        for (int i = 0; i <= nz - 1; ++i) {
            bool wall = true;
            if (i < nz - 1) {
                const int t = material_type[cell_material[(size_t)i + (size_t)j * (nz - 1)] + ndiel];
                wall = t == kMatMetal || t == kMatDiel;
            }
            if (wall) {
                if (i - run0 > best1 - best0) { best0 = run0; best1 = i; }
                run0 = i + 1;
            }
        }
        rv[(size_t)j] = int2{best0, best1};
    }
    HIPCHK(hipMalloc(&c->row_vac, sizeof(int2) * rv.size()));
    HIPCHK(hipMemcpy(c->row_vac, rv.data(), sizeof(int2) * rv.size(), hipMemcpyHostToDevice));
    c->h.row_vac = c->row_vac;
    c->h.cell_material = c->cell_material;
    c->h.hedge = c->hedge;
    c->h.vedge = c->vedge;
    c->h.mat_type = c->mat_type;
    c->h.ndiel = ndiel;
    c->h.nmetal = nmetal;
    c->h.metal_charge = c->metal_charge;
    c->geometry_set = true;
    return sync_state(c);
}

int ipm_upload_species(ipm_ctx *c, int s, int64_t n, const double *po, int layout)
{
    if (!c || s < 0 || s >= c->ns || !c->species_set[s]) return fail(IPM_EINVAL, "invalid species %d", s);
    if (n < 0 || n > c->h.cap[s]) return fail(IPM_ECAPACITY, "upload of %lld particles exceeds capacity %lld",
                                               (long long)n, (long long)c->h.cap[s]);
    if (n > 0 && !po) return fail(IPM_EINVAL, "null particle array");
    if (layout != IPM_LAYOUT_AOS && layout != IPM_LAYOUT_SOA) return fail(IPM_EINVAL, "invalid layout");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n > 0) {
        if (layout == IPM_LAYOUT_SOA) {
            for (int f = 0; f < 8; ++f)
                HIPCHK(hipMemcpy(c->h.f[s][f], po + (size_t)f * n, sizeof(double) * n, hipMemcpyHostToDevice));
        } else {
            std::vector<double> tmp((size_t)n);
            for (int f = 0; f < 8; ++f) {
                for (int64_t i = 0; i < n; ++i) tmp[i] = po[(size_t)i * 8 + f];
                HIPCHK(hipMemcpy(c->h.f[s][f], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
            }
        }
    }
    DevCounters ctr;
    HIPCHK(hipMemcpy(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost));
    ctr.n[s] = (unsigned long long)n; ctr.n_new[s] = 0; ctr.n_dead[s] = 0;
    ctr.overflow = 0;     // a fresh population: earlier overflow reports are void
    for (int k = 0; k < IPM_NS_MAX; ++k) ctr.q_total[k] = 0;
    HIPCHK(hipMemcpy(c->d_ctr, &ctr, sizeof ctr, hipMemcpyHostToDevice));
    return IPM_OK;
}

int ipm_upload_species_device(ipm_ctx *c, int s, int64_t offset, int64_t n, const double *po_dev, int64_t ld)
{
    if (!c || s < 0 || s >= c->ns || !c->species_set[s]) return fail(IPM_EINVAL, "invalid species %d", s);
    if (offset < 0 || n < 0 || ld < n) return fail(IPM_EINVAL, "invalid offset %lld / count %lld / leading dimension %lld",
                                                 (long long)offset, (long long)n, (long long)ld);
    if (offset + n > c->h.cap[s])
        return fail(IPM_ECAPACITY, "upload of slots [%lld, %lld) exceeds capacity %lld", (long long)offset,
                    (long long)(offset + n), (long long)c->h.cap[s]);
    if (n > 0 && !po_dev) return fail(IPM_EINVAL, "null particle array");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int f = 0; f < 8 && n > 0; ++f)
        HIPCHK(hipMemcpyAsync(c->h.f[s][f] + offset, po_dev + (size_t)f * (size_t)ld, sizeof(double) * (size_t)n,
                              hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    DevCounters ctr;
    HIPCHK(hipMemcpy(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost));
    ctr.n[s] = (unsigned long long)(offset + n); ctr.n_new[s] = 0; ctr.n_dead[s] = 0;
    ctr.overflow = 0;
    for (int k = 0; k < IPM_NS_MAX; ++k) ctr.q_total[k] = 0;
    HIPCHK(hipMemcpy(c->d_ctr, &ctr, sizeof ctr, hipMemcpyHostToDevice));
    return IPM_OK;
}

int ipm_download_species(ipm_ctx *c, int s, int64_t *n_out, double *po, int64_t cap, int layout)
{
    if (!c || s < 0 || s >= c->ns || !c->species_set[s] || !n_out) return fail(IPM_EINVAL, "invalid species %d", s);
    if (layout != IPM_LAYOUT_AOS && layout != IPM_LAYOUT_SOA) return fail(IPM_EINVAL, "invalid layout");
    DevCounters ctr;
    int rc = read_counters(c, &ctr);
    if (rc) return rc;
    const size_t n = (size_t)ctr.n[s];
    const size_t live = n - (size_t)ctr.n_dead[s];
    *n_out = (int64_t)live;
    if ((int64_t)live > cap) return fail(IPM_ECAPACITY, "download buffer holds %lld < %zu particles", (long long)cap, live);
    if (live == 0) return IPM_OK;
    if (!po) return fail(IPM_EINVAL, "null particle array");
    std::vector<double> tmp(n * 8);
    for (int f = 0; f < 8; ++f)
        HIPCHK(hipMemcpy(tmp.data() + f * n, c->h.f[s][f], sizeof(double) * n, hipMemcpyDeviceToHost));
    size_t j = 0;
    for (size_t i = 0; i < n; ++i) {
        if (tmp[i] < -1.0e299) continue;
        if (j >= live) return fail(IPM_EINVAL, "live count mismatch on species %d", s);
        for (int f = 0; f < 8; ++f) {
            double v = tmp[f * n + i];
            if (layout == IPM_LAYOUT_SOA) po[(size_t)f * live + j] = v;
            else po[j * 8 + f] = v;
        }
        ++j;
    }
    if (j != live) return fail(IPM_EINVAL, "live count mismatch on species %d (%zu vs %zu)", s, j, live);
    return IPM_OK;
}

int ipm_count(ipm_ctx *c, int s, int64_t *n)
{
    if (!c || s < 0 || s >= c->ns || !n) return fail(IPM_EINVAL, "invalid species %d", s);
    DevCounters ctr;
    int rc = read_counters(c, &ctr);
    if (rc) return rc;
    *n = (int64_t)(ctr.n[s] - ctr.n_dead[s]);
    return IPM_OK;
}

int ipm_set_fields(ipm_ctx *c, const double *Ez, const double *Er)
{
    if (!c || !Ez || !Er) return fail(IPM_EINVAL, "null field array");
    HIPCHK(hipMemcpyAsync(c->Ez, Ez, sizeof(double) * (c->nz + 1) * c->nr, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->Er, Er, sizeof(double) * c->nz * (c->nr + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->fields_set = true;
    return IPM_OK;
}

int ipm_set_potential(ipm_ctx *c, const double *phi)
{
    if (!c || !phi) return fail(IPM_EINVAL, "null argument");
    const size_t nn = (size_t)c->nz * c->nr;
    if (!c->phi) HIPCHK(hipMalloc(&c->phi, sizeof(double) * nn));
    HIPCHK(hipMemcpyAsync(c->phi, phi, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
    const long long tot = (long long)(c->nz + 1) * c->nr + (long long)c->nz * (c->nr + 1);
    hipLaunchKernelGGL(set_e_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, c->phi, c->Ez,
                       c->Er, c->nz, c->nr, c->dx);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->fields_set = true;
    return IPM_OK;
}

int ipm_set_solver_geometry(ipm_ctx *c, const double *cell_epsilon, const int *point_type)
{
    if (!c || !cell_epsilon || !point_type) return fail(IPM_EINVAL, "null argument");
    if (c->nz < 3 || c->nr < 3) return fail(IPM_EINVAL, "the field-solver system needs Nz, Nr >= 3");
    const size_t nn = (size_t)c->nz * c->nr, ncell = (size_t)(c->nz - 1) * (c->nr - 1);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->eps_cell) HIPCHK(hipMalloc(&c->eps_cell, sizeof(double) * ncell));
    if (!c->ptype) HIPCHK(hipMalloc(&c->ptype, sizeof(int) * nn));
    HIPCHK(hipMemcpy(c->eps_cell, cell_epsilon, sizeof(double) * ncell, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->ptype, point_type, sizeof(int) * nn, hipMemcpyHostToDevice));
    return IPM_OK;
}

int ipm_solver_system(ipm_ctx *c, int with_chi, const double *sigma, const double *bound, const int *bound_mask,
                      double *coe, double *source)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (!c->eps_cell || !c->ptype) return fail(IPM_EINVAL, "ipm_set_solver_geometry has not been called");
    if (!c->geometry_set) return fail(IPM_EINVAL, "ipm_set_geometry has not been called");
    if ((bound == nullptr) != (bound_mask == nullptr))
        return fail(IPM_EINVAL, "bound and bound_mask are given together or not at all");
    const size_t nn = (size_t)c->nz * c->nr;
    if (!c->sol) HIPCHK(hipMalloc(&c->sol, sizeof(double) * 8 * nn + sizeof(int) * nn));
    double *d_sigma = c->sol + 6 * nn, *d_bound = c->sol + 7 * nn;
    int *d_mask = reinterpret_cast<int *>(c->sol + 8 * nn);
    if (sigma) HIPCHK(hipMemcpyAsync(d_sigma, sigma, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
    if (bound) {
        HIPCHK(hipMemcpyAsync(d_bound, bound, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(d_mask, bound_mask, sizeof(int) * nn, hipMemcpyHostToDevice, c->stream));
    }
    // FS%Source = -1.d0 * dx * dx / Epsilon * (...): Fortran's unary minus
    // applies to the whole product, which rounds like this form
    const double scale = -(c->dx * c->dx / kEPS0);
    hipLaunchKernelGGL(solver_system_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, c->chi_tot,
                       c->eps_cell, c->ptype, c->rho_tot, sigma ? d_sigma : nullptr, c->inv_vol,
                       bound ? d_bound : nullptr, bound ? d_mask : nullptr, with_chi ? 1 : 0, scale, c->sol, c->nz,
                       c->nr);
    HIPCHK(hipGetLastError());
    if (coe) HIPCHK(hipMemcpyAsync(coe, c->sol, sizeof(double) * 5 * nn, hipMemcpyDeviceToHost, c->stream));
    if (source) HIPCHK(hipMemcpyAsync(source, c->sol + 5 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IPM_OK;
}

int ipm_split_merge(ipm_ctx *c, int s, int mode, double weight, double alpha, double grid_r, double merge_limit,
                    int64_t *stats)
{
    int rc = check_ready(c, false, false);
    if (rc) return rc;
    if (s < 0 || s >= c->ns) return fail(IPM_EINVAL, "invalid species index %d", s);
    if (mode < 0 || mode > 3) return fail(IPM_EINVAL, "invalid split/merge mode %d", mode);
    if (!(weight > 0.0) || !(grid_r > 0.0) || !std::isfinite(weight) || !std::isfinite(grid_r) ||
        !std::isfinite(alpha) || !std::isfinite(merge_limit))
        return fail(IPM_EINVAL, "split/merge needs finite weight > 0, grid_r > 0, alpha and merge_limit");
    if (!c->rezone) HIPCHK(hipMalloc(&c->rezone, sizeof(unsigned long long) * 3));
    HIPCHK(hipMemsetAsync(c->rezone, 0, sizeof(unsigned long long) * 3, c->stream));
    const long long cap = c->h.cap[s];
    if (cap > 0 && mode) {
        hipLaunchKernelGGL(split_merge_kernel, dim3((unsigned)((cap + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, c->d_state, c->d_ctr, s, mode, weight, alpha, grid_r, merge_limit, c->rezone);
        hipLaunchKernelGGL(rezone_fold_kernel, dim3(1), dim3(64), 0, c->stream, c->d_ctr, s, cap);
        HIPCHK(hipGetLastError());
    }
    unsigned long long t[3];
    HIPCHK(hipMemcpyAsync(t, c->rezone, sizeof t, hipMemcpyDeviceToHost, c->stream));
    DevCounters ctr;
    if ((rc = read_counters(c, &ctr))) return rc;
    if (stats)
        for (int q = 0; q < 3; ++q) stats[q] = (int64_t)t[q];
    return IPM_OK;
}

int ipm_set_step_rezone(ipm_ctx *c, int s, int mode, double weight, double alpha, double grid_r, double merge_limit)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (s < 0 || s >= c->ns) return fail(IPM_EINVAL, "invalid species index %d", s);
    if (mode < 0 || mode > 3) return fail(IPM_EINVAL, "invalid split/merge mode %d", mode);
    if (mode && (!(weight > 0.0) || !(grid_r > 0.0) || !std::isfinite(weight) || !std::isfinite(grid_r) ||
                 !std::isfinite(alpha) || !std::isfinite(merge_limit)))
        return fail(IPM_EINVAL, "split/merge needs finite weight > 0, grid_r > 0, alpha and merge_limit");
    bool before = false, after = false;
    for (int q = 0; q < c->ns; ++q) before = before || c->h.rz_mode[q] != 0;
    c->h.rz_mode[s] = mode;
    c->h.rz_weight[s] = weight;
    c->h.rz_alpha[s] = alpha;
    c->h.rz_gridr[s] = grid_r;
    c->h.rz_limit[s] = merge_limit;
    for (int q = 0; q < c->ns; ++q) after = after || c->h.rz_mode[q] != 0;
    if (before != after) drop_graphs(c);   // the step launches the other particle-pass variant
    return sync_state(c);
}

static int check_sel(ipm_ctx *c, int s)
{
    if (s < -1 || s >= c->ns) return fail(IPM_EINVAL, "invalid species index %d", s);
    return IPM_OK;
}

int ipm_move_es(ipm_ctx *c, int s)
{
    int rc = check_ready(c, true, false);
    if (rc || (rc = check_sel(c, s))) return rc;
    if ((rc = launch_particles<M_PUSH>(c, s))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return IPM_OK;
}

int ipm_absorb(ipm_ctx *c, int s)
{
    int rc = check_ready(c, false, true);
    if (rc || (rc = check_sel(c, s))) return rc;
    if ((rc = launch_particles<M_ABSORB>(c, s))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return IPM_OK;
}

int ipm_move_absorb(ipm_ctx *c, int s)
{
    int rc = check_ready(c, true, true);
    if (rc || (rc = check_sel(c, s))) return rc;
    if ((rc = launch_particles<M_PUSH | M_ABSORB>(c, s))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return IPM_OK;
}

int ipm_take_tallies(ipm_ctx *c, double *metal_charge, double *sigma)
{
    if (!c || !c->geometry_set) return fail(IPM_EINVAL, "geometry has not been set");
    const size_t nn = (size_t)c->nz * c->nr;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (metal_charge && c->h.nmetal > 0) {
        std::vector<double> t(c->h.nmetal);
        HIPCHK(hipMemcpy(t.data(), c->metal_charge, sizeof(double) * c->h.nmetal, hipMemcpyDeviceToHost));
        for (int m = 0; m < c->h.nmetal; ++m) metal_charge[m] += t[m];
    }
    if (sigma) {
        std::vector<double> t(nn);
        HIPCHK(hipMemcpy(t.data(), c->sigma, sizeof(double) * nn, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nn; ++k) sigma[k] += t[k];
    }
    hipLaunchKernelGGL(clear_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, c->sigma, nn);
    hipLaunchKernelGGL(clear_kernel, dim3((unsigned)((c->h.nmetal + 256) / 256)), dim3(256), 0, c->stream,
                       c->metal_charge, (size_t)(c->h.nmetal + 1));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return IPM_OK;
}

int ipm_p2c(ipm_ctx *c, int s, double *rho_one, double *chi_one)
{
    int rc = check_ready(c, false, true);
    if (rc || (rc = check_sel(c, s))) return rc;
    if ((rc = launch_particles<M_DEPOSIT>(c, s))) return rc;
    const int s0 = s < 0 ? 0 : s, s1 = s < 0 ? c->ns : s + 1;
    if ((rc = launch_grid(c, s0, s1, 0))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t nn = (size_t)c->nz * c->nr;
    if (rho_one) HIPCHK(hipMemcpy(rho_one, c->rho_one + (size_t)s0 * nn, sizeof(double) * nn * (s1 - s0), hipMemcpyDeviceToHost));
    if (chi_one) HIPCHK(hipMemcpy(chi_one, c->chi_one + (size_t)s0 * nn, sizeof(double) * nn * (s1 - s0), hipMemcpyDeviceToHost));
    return IPM_OK;
}

int ipm_moments(ipm_ctx *c, int s, int implicit_j, double *rho, double *energy_ev, double *mp, double *jz,
                double *jr)
{
    int rc = check_ready(c, implicit_j != 0, true);
    if (rc) return rc;
    if (s < 0 || s >= c->ns) return fail(IPM_EINVAL, "invalid species index %d", s);
    const size_t nn = (size_t)c->nz * c->nr;
    if (!c->mom) HIPCHK(hipMalloc(&c->mom, sizeof(double) * 5 * nn));
    HIPCHK(hipMemsetAsync(c->mom, 0, sizeof(double) * 5 * nn, c->stream));
    const long long cap = c->h.cap[s];
    if (cap > 0)
        hipLaunchKernelGGL(moments_kernel, dim3((unsigned)((cap + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                           c->d_state, c->d_ctr, s, implicit_j ? 1 : 0, c->mom);
    const double vf = c->h.vfactor;
    const double ef = 0.5 * c->h.mass[s] * vf * vf;
    hipLaunchKernelGGL(moments_scale_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, c->mom,
                       c->inv_vol, nn, ef, c->h.charge[s], vf);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    double *outs[5] = {rho, energy_ev, mp, jz, jr};
    for (int q = 0; q < 5; ++q)
        if (outs[q]) HIPCHK(hipMemcpy(outs[q], c->mom + (size_t)q * nn, sizeof(double) * nn, hipMemcpyDeviceToHost));
    return IPM_OK;
}

int ipm_moments_noalloc(ipm_ctx *c, int s, int implicit_j, double *rho, double *energy_ev, double *mp, double *jz,
                        double *jr)
{
    int rc = check_ready(c, implicit_j != 0, true);
    if (rc) return rc;
    if (s < 0 || s >= c->ns) return fail(IPM_EINVAL, "invalid species index %d", s);
    const size_t nn = (size_t)c->nz * c->nr;
    if (nn > kScratchNodes)
        return fail(IPM_EINVAL, "grid of %zu nodes exceeds the %zu nodes of the no-allocation scratch", nn,
                    kScratchNodes);
    if (!c->scratch) return fail(IPM_EINVAL, "no scratch address for device %d", c->device);
    double *acc = c->scratch;
    // the accumulators keep the previous call's contents: clear them with a
    // kernel (no memset node, so a captured graph replays it correctly)
    hipLaunchKernelGGL(clear_kernel, dim3((unsigned)((5 * nn + 255) / 256)), dim3(256), 0, c->stream, acc, 5 * nn);
    const long long cap = c->h.cap[s];
    if (cap > 0)
        hipLaunchKernelGGL(moments_kernel, dim3((unsigned)((cap + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                           c->d_state, c->d_ctr, s, implicit_j ? 1 : 0, acc);
    const double vf = c->h.vfactor;
    const double ef = 0.5 * c->h.mass[s] * vf * vf;
    hipLaunchKernelGGL(moments_scale_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, acc,
                       c->inv_vol, nn, ef, c->h.charge[s], vf);
    HIPCHK(hipGetLastError());
    double *outs[5] = {rho, energy_ev, mp, jz, jr};
    for (int q = 0; q < 5; ++q)
        if (outs[q])
            HIPCHK(hipMemcpyAsync(outs[q], acc + (size_t)q * nn, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                  c->stream));
    return IPM_OK;
}

static int field_sum_device(ipm_ctx *c)
{
    return launch_grid(c, 0, 0, 1);
}

int ipm_field_sum(ipm_ctx *c, double *rho, double *chi)
{
    int rc = check_ready(c, false, false);
    if (rc) return rc;
    if ((rc = field_sum_device(c))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (rho) HIPCHK(hipMemcpy(rho, c->rho_tot, sizeof(double) * c->nz * c->nr, hipMemcpyDeviceToHost));
    if (chi) HIPCHK(hipMemcpy(chi, c->chi_tot, sizeof(double) * (c->nz + 2) * (c->nr + 2), hipMemcpyDeviceToHost));
    return IPM_OK;
}

int ipm_set_mcc(ipm_ctx *c, int s, int g, int model, int nreaction, int nsigma, double emin, double einterval,
                double emax, double collision_ratio, const int *rtype, const int *resultant, const double *threshold,
                const double *prob, double gas_mass, double gas_temperature, double beta_max)
{
    if (!c || s < 0 || s >= c->ns || g < 0 || g >= c->ng) return fail(IPM_EINVAL, "invalid (species, gas) = (%d, %d)", s, g);
    if (model < 0 || model > 3) return fail(IPM_EINVAL, "invalid MCC model %d", model);
    DevTable &t = c->h.tab[s * c->ng + g];
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void *p : c->tab_mem[s * c->ng + g]) HIPCHK(hipFree(p));
    c->tab_mem[s * c->ng + g].clear();
    memset(&t, 0, sizeof t);
    t.model = model;
    if (model != 0) {
        if (nreaction < 1 || nreaction > IPM_NR_MAX || nsigma < 1 || !rtype || !resultant || !threshold || !prob ||
            !(gas_mass > 0) || !(gas_temperature > 0) || !(einterval > 0))
            return fail(IPM_EINVAL, "invalid MCC table for (species, gas) = (%d, %d)", s, g);
        if (model == 3 && nreaction < 2) return fail(IPM_EINVAL, "reactive table needs >= 2 reactions");
        for (int j = 0; j < nreaction; ++j)
            if (resultant[j] < -1 || resultant[j] >= 100)
                return fail(IPM_EINVAL, "invalid resultant %d", resultant[j]);
        int *d_rt, *d_res;
        double *d_th, *d_p;
        HIPCHK(hipMalloc(&d_rt, sizeof(int) * nreaction));
        HIPCHK(hipMalloc(&d_res, sizeof(int) * nreaction));
        HIPCHK(hipMalloc(&d_th, sizeof(double) * nreaction));
        HIPCHK(hipMalloc(&d_p, sizeof(double) * nreaction * nsigma));
        c->tab_mem[s * c->ng + g] = {d_rt, d_res, d_th, d_p};
        HIPCHK(hipMemcpy(d_rt, rtype, sizeof(int) * nreaction, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_res, resultant, sizeof(int) * nreaction, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_th, threshold, sizeof(double) * nreaction, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_p, prob, sizeof(double) * nreaction * nsigma, hipMemcpyHostToDevice));
        t.nreaction = nreaction; t.nsigma = nsigma;
        t.emin = emin; t.einterval = einterval; t.emax = emax; t.ratio = collision_ratio;
        t.gas_mass = gas_mass; t.gas_temperature = gas_temperature; t.beta_max = beta_max;
        t.rtype = d_rt; t.resultant = d_res; t.threshold = d_th; t.prob = d_p;
        bool cum = true;
        for (int i = 0; i < nsigma && cum; ++i)
            for (int j = 0; j < nreaction; ++j) {
                const double v = prob[(size_t)i * nreaction + j];
                if (!(v >= 0.0) || (j > 0 && v < prob[(size_t)i * nreaction + j - 1])) { cum = false; break; }
            }
        t.cumulative = cum ? 1 : 0;
        if (model == 1 || model == 2) {
            const std::vector<double> nb = null_bound_table(nreaction, nsigma, prob);
            if (!nb.empty()) {
                double *d_nb;
                HIPCHK(hipMalloc(&d_nb, sizeof(double) * nb.size()));
                c->tab_mem[s * c->ng + g].push_back(d_nb);
                HIPCHK(hipMemcpy(d_nb, nb.data(), sizeof(double) * nb.size(), hipMemcpyHostToDevice));
                t.nb_tab = d_nb;
                t.nb_eps = 0.01 + 4e-6 / einterval;
                t.nb_inv_einterval = 1.0 / einterval;
                t.nb_inv_emax = 1.0 / emax;
            }
        }
    }
    table_consts(c);
    int rc = ensure_queue(c);
    if (rc) return rc;
    return sync_state(c);
}

int ipm_set_rng(ipm_ctx *c, uint64_t seed, uint32_t step)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    DevCounters ctr;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost));
    ctr.seed = seed;
    ctr.step = step;
    ctr.overflow = 0;     // a restart of the stream clears earlier overflow reports
    for (int k = 0; k < IPM_NS_MAX; ++k) ctr.q_total[k] = 0;
    HIPCHK(hipMemcpy(c->d_ctr, &ctr, sizeof ctr, hipMemcpyHostToDevice));
    return IPM_OK;
}

int ipm_mcc(ipm_ctx *c)
{
    int rc = check_ready(c, false, false);
    if (rc) return rc;
    if ((rc = launch_particles<M_MCC>(c, -1))) return rc;
    if ((rc = launch_collide(c, true))) return rc;
    DevCounters ctr;
    return read_counters(c, &ctr);
}

int ipm_set_event_counting(ipm_ctx *c, int on)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    c->h.count_events = on ? 1 : 0;
    return sync_state(c);
}

int ipm_set_null_screen(ipm_ctx *c, int on)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    c->h.null_screen = on ? 1 : 0;
    return sync_state(c);
}

int ipm_take_mcc_counts(ipm_ctx *c, int64_t *counts)
{
    if (!c || !counts) return fail(IPM_EINVAL, "null argument");
    const size_t n = (size_t)c->ns * c->ng * 128;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(counts, c->mcc_counts, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemset(c->mcc_counts, 0, sizeof(unsigned long long) * IPM_NS_MAX * IPM_NG_MAX * 128));
    return IPM_OK;
}

static int timed_event(ipm_ctx *c, int *idx)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    *idx = (int)c->ev_used++;
    HIPCHK(hipEventRecord(c->ev_pool[*idx], c->stream));
    return IPM_OK;
}

// Enqueues one step.  With timing on (and not capturing), each of the three
// kernels is bracketed by HIP events on the library's stream.
static int step_enqueue(ipm_ctx *c, bool capturing)
{
    const bool t = c->timing && !capturing;
    int a = 0, b = 0, rc;
    if (t && (rc = timed_event(c, &a))) return rc;
    bool rezone = false;
    for (int q = 0; q < c->ns; ++q) rezone = rezone || c->h.rz_mode[q] != 0;
    if (rezone) rc = launch_particles<M_PUSH | M_ABSORB | M_DEPOSIT | M_MCC | M_REZONE>(c, -1);
    else rc = launch_particles<M_PUSH | M_ABSORB | M_DEPOSIT | M_MCC>(c, -1);
    if (rc) return rc;
    if (t) {
        if ((rc = timed_event(c, &b))) return rc;
        c->ev_pairs[0].push_back({a, b});
        a = b;
    }
    if ((rc = launch_collide(c, false))) return rc;
    if (t) {
        if ((rc = timed_event(c, &b))) return rc;
        c->ev_pairs[1].push_back({a, b});
        a = b;
    }
    if ((rc = launch_grid(c, 0, c->ns, 1, 1))) return rc;
    if (t) {
        if ((rc = timed_event(c, &b))) return rc;
        c->ev_pairs[2].push_back({a, b});
    }
    return IPM_OK;
}

int ipm_step(ipm_ctx *c)
{
    int rc = check_ready(c, true, true);
    if (rc || (rc = step_enqueue(c, false))) return rc;
    if (c->sort_every > 0 && ++c->since_sort >= c->sort_every) {
        if ((rc = sort_enqueue(c, 0, c->ns))) return rc;
        sort_mirror(c, 0, c->ns);
        c->since_sort = 0;
    }
    return IPM_OK;
}

int ipm_set_timing(ipm_ctx *c, int on)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->timing = on != 0;
    for (auto &v : c->ev_pairs) v.clear();
    c->ev_used = 0;
    return IPM_OK;
}

int ipm_take_timing(ipm_ctx *c, double *ms, int64_t *launches)
{
    if (!c || !ms || !launches) return fail(IPM_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) {
        double tot = 0.0;
        for (auto &pr : c->ev_pairs[k]) {
            float e = 0.0f;
            HIPCHK(hipEventElapsedTime(&e, c->ev_pool[pr.first], c->ev_pool[pr.second]));
            tot += e;
        }
        ms[k] = tot;
        launches[k] = (int64_t)c->ev_pairs[k].size();
        c->ev_pairs[k].clear();
    }
    c->ev_used = 0;
    return IPM_OK;
}

int ipm_step_graph(ipm_ctx *c)
{
    int rc = check_ready(c, true, true);
    if (rc) return rc;
    if (!c->graph) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = step_enqueue(c, true);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return fail(IPM_EHIP, "graph capture failed: %s", hipGetErrorString(e));
        HIPCHK(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
    }
    HIPCHK(hipGraphLaunch(c->graph, c->stream));
    if (c->sort_every > 0 && ++c->since_sort >= c->sort_every) {
        if ((rc = sort_graph_launch(c))) return rc;
        c->since_sort = 0;
    }
    return IPM_OK;
}

int ipm_synchronize(ipm_ctx *c)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    DevCounters ctr;
    return read_counters(c, &ctr);
}

int ipm_sort(ipm_ctx *c, int s)
{
    int rc = check_ready(c, false, false);
    if (rc || (rc = check_sel(c, s))) return rc;
    const int s0 = s < 0 ? 0 : s, s1 = s < 0 ? c->ns : s + 1;
    DevCounters ctr;
    if ((rc = read_counters(c, &ctr))) return rc;
    if ((rc = sort_enqueue(c, s0, s1))) return rc;
    sort_mirror(c, s0, s1);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->since_sort = 0;
    return IPM_OK;
}

int ipm_set_sort_interval(ipm_ctx *c, int steps)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (steps < 0) return fail(IPM_EINVAL, "sort interval must be >= 0, got %d", steps);
    c->sort_every = steps;
    c->since_sort = 0;
    return IPM_OK;
}

int ipm_set_sort_tile(ipm_ctx *c, int slots)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (slots != kSortTile && slots != kSortTileMax)
        return fail(IPM_EINVAL, "sort tile must be %d or %d slots, got %d", kSortTile, kSortTileMax, slots);
    c->sort_block = slots / kSortItems;
    drop_graphs(c);
    return IPM_OK;
}

int ipm_set_launch(ipm_ctx *c, int block, int blocks_per_cu)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (block != 0 && block != 128 && block != 256 && block != 512)
        return fail(IPM_EINVAL, "block must be 0 (256), 128, 256 or 512 threads, got %d", block);
    if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(IPM_EINVAL, "invalid blocks_per_cu %d", blocks_per_cu);
    const int b = block ? block : kBlock;
    if (b != kBlock && c->pk_winr != kWinR)
        return fail(IPM_EINVAL, "a %d-thread block runs with the %d-row window only", b, kWinR);
    c->pk_block = b;
    c->blocks_per_cu = blocks_per_cu;
    drop_graphs(c);
    return IPM_OK;
}

int ipm_set_deposit_window(ipm_ctx *c, int rows)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (rows != 16 && rows != 24 && rows != 32) return fail(IPM_EINVAL, "window rows must be 16, 24 or 32, got %d", rows);
    if (rows != kWinR && c->pk_block != kBlock)
        return fail(IPM_EINVAL, "a %d-row window runs with %d-thread blocks only", rows, kBlock);
    c->pk_winr = rows;
    drop_graphs(c);
    return IPM_OK;
}

int ipm_set_collide_grid(ipm_ctx *c, int blocks_per_cu)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (blocks_per_cu < 1 || blocks_per_cu > kCollideBlocksPerCuMax)
        return fail(IPM_EINVAL, "collision blocks per CU must be 1..%d, got %d", kCollideBlocksPerCuMax, blocks_per_cu);
    c->collide_bpc = blocks_per_cu;
    drop_graphs(c);
    if (c->num_cu * blocks_per_cu < c->ns)
        return fail(IPM_EINVAL, "collision grid of %d blocks is smaller than the %d species", c->num_cu * blocks_per_cu,
                    c->ns);
    if (!c->p_stage) return IPM_OK;       // applied when the collision buffers are made
    c->h.c_grid = c->num_cu * c->collide_bpc;
    return sync_state(c);
}

int ipm_species_device_ptrs(ipm_ctx *c, int s, double **fields8)
{
    if (!c || s < 0 || s >= c->ns || !fields8 || !c->species_set[s]) return fail(IPM_EINVAL, "invalid species %d", s);
    for (int f = 0; f < 8; ++f) fields8[f] = c->h.f[s][f];
    return IPM_OK;
}

}  // extern "C"
