// Device-side building blocks of the gfx950 particle hot path.
//
// Each routine is a restatement of the reference routine cited beside it,
// written so that every product and sum is rounded in the order the Fortran
// source evaluates it (the library is compiled with -ffp-contract=off, so no
// FMA contraction moves a result).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ipm_hip.h"

namespace ipm {

// R/src/base/Constants.F90:3-15
constexpr double kPI = 3.141592653589793238;
constexpr double kKB = 1.3807e-23;
constexpr double kEXP = 2.71828;
constexpr double kJTOEV = 1.6022e-19;
constexpr double kEPS0 = 8.8542e-12;
constexpr double kMINREAL = 1.0e-15;

constexpr int kMatMetal = 1;       // R/src/pic/materials/Materials.F90:10-12
constexpr int kMatDiel = 2;
constexpr int kEdgeV2D = 1;        // R/src/base/Geometry.F90:15-19
constexpr int kEdgeD2V = 3;

// Z of a deleted slot.  No push from a position inside the grid can reach
// it, so slots pushed to Z <= 0 and awaiting absorption stay distinguishable.
constexpr double kDead = -1.0e300;

struct DevTable {
    int model, nreaction, nsigma;
    int cumulative;         // every column of prob nondecreasing in j and >= 0 (checked on the host)
    double emin, einterval, emax, ratio;
    double gas_mass, gas_temperature, beta_max, pad2;
    // per-table constants of Updater and VelMaxInit, formed on the host by
    // the same expressions: miu = sm*gm/(sm+gm), mass_ratio = gm/(sm+gm),
    // beta = 1/(kB*T)
    double miu, mass_ratio, beta, pad3;
    const int *rtype;
    const int *resultant;
    const double *threshold;
    const double *prob;     // Probility(j,i) -> prob[(i-1)*nreaction + (j-1)]
    // Null-event screen (models 1 and 2; null: no screen).  nb_tab holds
    // levels L = 0, 1, ... of (nsigma + 1) doubles: entry b of level L bounds
    // every value SelectProbility can interpolate in energy bins b .. b+2^L-1
    // (bin 0: N < 1; bin b < nsigma: rows b-1 and b; bin nsigma: the last row,
    // before its E/emax factor).  nb_vmax bounds the speed VelMaxInit can
    // draw, nb_ecoef = miu/2/JTOEV, nb_eps the error of approx_lx in bins.
    const double *nb_tab;
    double nb_vmax, nb_ecoef, nb_eps, nb_inv_einterval, nb_inv_emax, pad4;
};

// Device-resident description of the whole problem: one pointer is the only
// kernel argument, so a captured graph keeps working after buffers swap.
struct DevState {
    int nz, nr, ns, ng;
    double dx, dt, vfactor, inv_vfactor;
    double *f[IPM_NS_MAX][8];          // Z R Vz Vr Vt Az Ar WQ
    double *alt[IPM_NS_MAX][8];        // the other buffer set: the cell sort's target
    long long cap[IPM_NS_MAX];
    double charge[IPM_NS_MAX], mass[IPM_NS_MAX], efactor[IPM_NS_MAX];
    double *raw[IPM_NS_MAX];           // raw per-species deposit (nz*nr)
    const double *Ez, *Er;
    const int *cell_material, *hedge, *vedge, *mat_type;
    // Per cell row j (0-based, nr + 2 entries, the last three empty): a run [x, y) of
    // 0-based z cell indices whose cells are neither metal nor dielectric,
    // so absorption needs no material lookup there (formed on the host from
    // cell_material and mat_type).
    const int2 *row_vac;
    int ndiel, nmetal;
    double *metal_charge;              // nmetal
    double *sigma;                     // nz*nr
    unsigned long long *mcc_counts;    // optional: (s*ng+g)*128 + j
    int count_events;
    int null_screen;                   // use the tables' null-event screen
    // Queue of particles selected for a collision (null-collision pass).
    // Species s owns q_maxtiles[s] = ceil(cap/kTile) tile regions of
    // q_tile_cap[s] entries; entry k of live tile lt of species s sits at
    // q_region[s] + lt*q_tile_cap[s] + k and holds the particle's state and
    // its index inside the tile.  q_tile_count[q_cbase[s] + lt] is that
    // tile's number of entries in the current pass.  The layout needs no
    // global counter: a tile appends with an LDS counter only.  Entries are
    // 64-byte records (Z R Vz Vr Vt WQ, index in tile, pad), written and read whole.
    double *q_rec;
    unsigned int *q_tile_count;
    // Per tile (indexed like q_tile_count): the (z, r) corner of its LDS
    // deposit window, set from the tile's first particle by the previous pass.
    int *tile_anchor;
    long long q_region[IPM_NS_MAX], q_cbase[IPM_NS_MAX], q_tile_cap[IPM_NS_MAX];
    double q_weight[IPM_NS_MAX];       // expected entries per tile (collision-pass split)
    // Per collision block (c_grid blocks): the products it staged
    // (kProdStage records of 8 doubles, each with species | rank << 8, and
    // per-species counts) for the append pass.
    double *p_stage;
    unsigned *p_meta;
    unsigned *p_cnt;
    int c_grid;
    int c_pad;
    DevTable tab[IPM_NS_MAX * IPM_NG_MAX];
    // Split/merge inside the fused step (ipm_set_step_rezone): per species
    // the mode (1 split, 2 merge, 0 none) and CalRW's Weight, alpha, grid_r
    // and the merge limit.
    int rz_mode[IPM_NS_MAX];
    double rz_weight[IPM_NS_MAX], rz_alpha[IPM_NS_MAX], rz_gridr[IPM_NS_MAX], rz_limit[IPM_NS_MAX];
};

// Mutable per-step counters (device memory).
struct DevCounters {
    unsigned long long n[IPM_NS_MAX];      // slots in use at step start
    unsigned long long n_new[IPM_NS_MAX];  // products appended during the pass
    unsigned long long n_dead[IPM_NS_MAX]; // dead slots among n
    unsigned long long q_total[IPM_NS_MAX]; // collision-queue entries of the pass (cleared by the append)
    unsigned long long seed;
    unsigned int step;
    unsigned int blocks_done;
    int overflow;                          // 1 capacity, 2 bad species, 4 queue, 8 dead-slot count
    int pad;
};

// Global-address-space views of device pointers: loads, stores and atomics
// through them are global_* instructions instead of flat_*.
// Element at a 32-bit byte offset from a base pointer: the address is the
// base (scalar registers) plus one zero-extended vector register, the SADDR
// form of the global instructions, with no 64-bit address arithmetic per
// access.  Particle arrays are limited to 2^32 bytes per field for it.
__device__ __forceinline__ __attribute__((address_space(1))) double *at_off(double *base, unsigned boff)
{
    return (__attribute__((address_space(1))) double *)((__attribute__((address_space(1))) char *)base + boff);
}
__device__ __forceinline__ const __attribute__((address_space(1))) double *at_off(const double *base, unsigned boff)
{
    return (const __attribute__((address_space(1))) double *)((const __attribute__((address_space(1))) char *)base + boff);
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gp(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gp(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}

__device__ __forceinline__ void global_add(double *a, double v)
{
    __hip_atomic_fetch_add(gp(a), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Device-scope atomics through global-address-space views (flat atomics
// would also count as LDS operations, so that every later wait for LDS
// waited for them too).
template <class T>
__device__ __forceinline__ T global_fetch_add(T *a, T v)
{
    return __hip_atomic_fetch_add(gp(a), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void global_or(int *a, int v)
{
    __hip_atomic_fetch_or(gp(a), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Pointers into LDS, typed so: accesses through them are ds_* instructions.
template <class T>
using lds_ptr = __attribute__((address_space(3))) T *;

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 multiply each (v_mad_u64_u32) instead of hi/lo pairs
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

// Per-particle random stream.  The counter carries 96 bits folded from the
// particle's phase-space state at the start of the MCC pass plus the block
// index; the key carries the seed, the step and the (species, gas) tag.  The
// stream therefore depends on what the particle is, not on where it is
// stored: sorting, compaction and the order in which products are appended
// do not change any draw.  One Philox block gives two draws: draw k is the
// top 52 bits of words (1,0) of block k/2 for even k and of words (3,2) for
// odd k, so a run of draws costs one block per two draws (the target
// sampler's rejection loop draws in pairs).
struct Rng {
    uint32_t c0, c1, c2, k;
    uint32_t key0, key1;
    uint64_t spare;     // draw k's 52 bits when k is odd and have is set
    bool have;
    __device__ __forceinline__ double next()
    {
        uint64_t x;
        if ((k & 1u) && have) {
            x = spare;
        } else {
            uint32_t c[4] = {c0, c1, c2, k >> 1};
            philox4x32_10(c, key0, key1);
            const uint64_t lo = ((((uint64_t)c[1]) << 32) | c[0]) >> 12;   // 52 bits
            const uint64_t hi = ((((uint64_t)c[3]) << 32) | c[2]) >> 12;
            x = (k & 1u) ? hi : lo;
            spare = hi;
            have = true;
        }
        ++k;
        return ((double)x + 0.5) * 0x1p-52;                     // (0,1)
    }
    // continue the stream at draw k (nothing of the current block is kept)
    __device__ __forceinline__ void seek(uint32_t draw)
    {
        k = draw;
        have = false;
    }
};

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

struct StateKey {
    uint32_t c0, c1, c2;
};

__device__ __forceinline__ StateKey state_key(double Z, double R, double Vz, double Vr, double Vt)
{
    const uint64_t z = (uint64_t)__double_as_longlong(Z), r = (uint64_t)__double_as_longlong(R);
    const uint64_t vz = (uint64_t)__double_as_longlong(Vz), vr = (uint64_t)__double_as_longlong(Vr);
    const uint64_t vt = (uint64_t)__double_as_longlong(Vt);
    StateKey s;
    s.c0 = (uint32_t)z ^ (uint32_t)(vr >> 32) ^ rotl32((uint32_t)(vt >> 32), 13);
    s.c1 = (uint32_t)r ^ (uint32_t)(vz >> 32) ^ rotl32((uint32_t)vt, 7);
    s.c2 = (uint32_t)vz ^ rotl32((uint32_t)vr, 11) ^ rotl32((uint32_t)(z >> 32), 3) ^ rotl32((uint32_t)(r >> 32), 19);
    return s;
}

__device__ __forceinline__ Rng make_rng(const StateKey &sk, unsigned long long seed, uint32_t step, uint32_t tag)
{
    Rng g;
    g.c0 = sk.c0; g.c1 = sk.c1; g.c2 = sk.c2; g.k = 0;
    g.spare = 0; g.have = false;
    g.key0 = (uint32_t)seed + step * 0x9E3779B9u;
    g.key1 = (uint32_t)(seed >> 32) ^ (tag * 0x85EBCA6Bu);
    return g;
}

// ---------------------------------------------------------------- push
struct Part {
    double Z, R, Vz, Vr, Vt, Az, Ar, WQ;
};

// Field sources for the gather of the push.  GridField reads one Ez/Er pair;
// PairField reads interleaved (E0, E1) pairs and returns E0 + c * E1,
// rounded as c * E1 first and the sum second (the same two roundings as
// forming the combined grid on the host and gathering from it).
struct GridField {
    const double *Ez, *Er;
    __device__ __forceinline__ double ez(size_t k) const { return gp(Ez)[k]; }
    __device__ __forceinline__ double er(size_t k) const { return gp(Er)[k]; }
};
struct alignas(16) Pair2 {
    double x, y;
};
// one 16-byte global load of pair k
__device__ __forceinline__ Pair2 ld_pair(const Pair2 *p, size_t k)
{
    const __attribute__((address_space(1))) Pair2 *q = gp(p) + k;
    Pair2 v;
    v.x = q->x;
    v.y = q->y;
    return v;
}
struct PairField {
    const Pair2 *Ez, *Er;
    double c;
    __device__ __forceinline__ double ez(size_t k) const { const Pair2 v = ld_pair(Ez, k); return v.x + c * v.y; }
    __device__ __forceinline__ double er(size_t k) const { const Pair2 v = ld_pair(Er, k); return v.x + c * v.y; }
};

// Gather indices and weights of MoveESIm, R/src/pic/particle/ParticleBundle.f90:685-700.
// Index guards keep every read inside the arrays for any input; for the
// positions absorption leaves behind (0 < Z < Nz-1, 0 <= R < Nr-1) they
// are no-ops except R == 0 exactly, whose N2 row carries weight R1 = 0.
struct Gather {
    double Z1, Z2, Z1H, Z2H, R1, R2, R1H, R2H;
    size_t kr, kz0, kz1, ldr;
};
__device__ __forceinline__ Gather gather_at(int nz, int nr, double z, double r)
{
    Gather g;
    const size_t ldz = (size_t)nz + 1;   // Ez(0:nz, 1:nr)
    g.ldr = (size_t)nz;                  // Er(1:nz, 0:nr)
    int N1 = (int)ceil(z);
    g.Z1 = (double)N1 - z; g.Z2 = 1.0 - g.Z1;
    int N1H = (int)(z + 0.5);
    g.Z1H = (double)N1H + 0.5 - z; g.Z2H = 1.0 - g.Z1H;
    int N2 = (int)ceil(r);
    g.R1 = (double)N2 - r; g.R2 = 1.0 - g.R1;
    int N2H = (int)(r + 0.5);
    g.R1H = (double)N2H + 0.5 - r; g.R2H = 1.0 - g.R1H;
    int i1 = min(max(N1, 1), nz - 1);
    int i1h = min(max(N1H, 0), nz - 1);
    int j2 = min(max(N2, 1), nr - 1);
    int j2b = min(max(N2, 0), nr - 1) + 1;
    int j2h = min(max(N2H, 0), nr - 1);
    g.kr = (size_t)(i1 - 1) + (size_t)j2h * g.ldr;
    g.kz0 = (size_t)i1h + (size_t)(j2 - 1) * ldz;
    g.kz1 = (size_t)i1h + (size_t)(j2b - 1) * ldz;
    return g;
}

// MoveESIm for one particle, R/src/pic/particle/ParticleBundle.f90:685-745.
template <class FS>
__device__ __forceinline__ void push_core(const FS &fs, int nz, int nr, double EFactor, Part &p)
{
    const double z = p.Z, r = p.R;
    const Gather g = gather_at(nz, nr, z, r);
    const size_t kr = g.kr, kz0 = g.kz0, kz1 = g.kz1, ldr = g.ldr;
    const double Z1 = g.Z1, Z2 = g.Z2, Z1H = g.Z1H, Z2H = g.Z2H, R1 = g.R1, R2 = g.R2, R1H = g.R1H, R2H = g.R2H;
    // the eight field values are read before any is used, so that their
    // loads are in flight together (one wait for all of them)
    const double er0 = fs.er(kr), er1 = fs.er(kr + 1), er2 = fs.er(kr + ldr), er3 = fs.er(kr + ldr + 1);
    const double ez0 = fs.ez(kz0), ez1 = fs.ez(kz0 + 1), ez2 = fs.ez(kz1), ez3 = fs.ez(kz1 + 1);
    __builtin_amdgcn_sched_barrier(0);
    double E1 = er0 * Z1 * R1H + er1 * Z2 * R1H + er2 * Z1 * R2H + er3 * Z2 * R2H;
    double E2 = ez0 * Z1H * R1 + ez1 * Z2H * R1 + ez2 * Z1H * R2 + ez3 * Z2H * R2;
    E1 = E1 * EFactor;
    E2 = E2 * EFactor;

    double vz = p.Vz + E2;
    double zz = z + E2;
    double az = (p.Az + E2) * 0.5;
    double vr = p.Vr + E1;
    double rr = r + E1;
    double ar = (p.Ar + E1) * 0.5;

    vz = vz + az;
    zz = zz + vz;

    double Vx = vr + ar;
    double Vy = p.Vt;
    double X = rr + Vx;
    double Y = Vy;
    double P_R = sqrt(X * X + Y * Y);
    double vt;
    if (P_R < kMINREAL) {
        vr = sqrt(Vx * Vx + Vy * Vy);
        vt = kMINREAL;
        ar = fabs(ar);
    } else {
        double SinA = Y / P_R;
        double CosA = X / P_R;
        vr = Vx * CosA + Vy * SinA;
        vt = -Vx * SinA + Vy * CosA;
        ar = ar * CosA;
    }
    p.Z = zz; p.R = P_R; p.Vz = vz; p.Vr = vr; p.Vt = vt; p.Az = az; p.Ar = ar;
}

__device__ __forceinline__ void push_one(const DevState &st, double EFactor, Part &p)
{
    push_core(GridField{st.Ez, st.Er}, st.nz, st.nr, EFactor, p);
}

// ---------------------------------------------------------- absorption
// MaterialAborptionParticle, R/src/pic/particleboundary/ParticleMaterial.F90:22-85,
// for a particle inside the domain (0 < Z < Nz-1, 0 < R < Nr-1) in cell
// (NzL, NrL): metal takes its charge, a dielectric cell puts it on the edge
// the reference chooses; either deletes the particle (returns true).
// metal_acc (LDS, kMetalLds entries, or null): per-block sums of the charge
// of metals 1..kMetalLds, added to metal_charge once per block by the
// caller (every absorbed particle of a metal would otherwise add to one
// address, and atomics on one address serialize).
constexpr int kMetalLds = 16;
__device__ __forceinline__ bool absorb_material(const DevState &st, double charge, const Part &p, int NzL, int NrL,
                                                double *metal_acc = nullptr)
{
    const int nz = st.nz;
    int m = gp(st.cell_material)[(size_t)(NzL - 1) + (size_t)(NrL - 1) * (nz - 1)];
    int t = gp(st.mat_type)[m + st.ndiel];
    if (t == kMatMetal) {
        if (metal_acc && m <= kMetalLds)
            __hip_atomic_fetch_add(&metal_acc[m - 1], p.WQ * charge, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            global_add(&st.metal_charge[m - 1], p.WQ * charge);
        return true;
    }
    if (t == kMatDiel) {
        const auto *he = gp(st.hedge), *ve = gp(st.vedge);
        double *sg = st.sigma;
        auto HE = [&](int i, int j) { return he[(size_t)(i - 1) + (size_t)(j - 1) * nz]; };
        auto VE = [&](int i, int j) { return ve[(size_t)(i - 1) + (size_t)(j - 1) * (nz - 1)]; };
        auto SG = [&](int i, int j) { return sg + (size_t)(i - 1) + (size_t)(j - 1) * nz; };
        int h0 = HE(NzL, NrL), h1 = HE(NzL + 1, NrL);
        int v0 = VE(NzL, NrL), v1 = VE(NzL, NrL + 1);
        if (h0 == kEdgeV2D || h0 == kEdgeD2V) {
            double RL = (double)NrL - p.R, RU = 1.0 - RL;
            global_add(SG(NzL, NrL), RL * p.WQ * charge);
            global_add(SG(NzL, NrL + 1), RU * p.WQ * charge);
        } else if (h1 == kEdgeV2D || h1 == kEdgeD2V) {
            double RL = (double)NrL - p.R, RU = 1.0 - RL;
            global_add(SG(NzL + 1, NrL), RL * p.WQ * charge);
            global_add(SG(NzL + 1, NrL + 1), RU * p.WQ * charge);
        } else if (v0 == kEdgeV2D || v0 == kEdgeD2V) {
            double ZL = (double)NzL - p.Z, ZU = 1.0 - ZL;
            global_add(SG(NzL, NrL), ZL * p.WQ * charge);
            global_add(SG(NzL + 1, NrL), ZU * p.WQ * charge);
        } else if (v1 == kEdgeV2D || v1 == kEdgeD2V) {
            double ZL = (double)NzL - p.Z, ZU = 1.0 - ZL;
            global_add(SG(NzL, NrL + 1), ZL * p.WQ * charge);
            global_add(SG(NzL + 1, NrL + 1), ZU * p.WQ * charge);
        }
        return true;
    }
    return false;
}

// Vacuum runs of cell rows r0 .. r0 + 3 as row_vac holds them (n = 0: none).
struct RowRuns {
    int r0 = 0, n = 0;
    int2 v0{0, 0}, v1{0, 0}, v2{0, 0}, v3{0, 0};
};

// MaterialAborptionParticle for one particle,
// R/src/pic/particleboundary/ParticleMaterial.F90:10-90: deleted outside
// the domain, reflected at the axis, else the material rules of the cell.
// Returns true if the particle is deleted; reflection edits p.R/Vr/Ar.  A
// particle whose cell lies in one of the vacuum runs `rr` is known to be in
// vacuum without the two dependent material loads.  Only the material
// lookup is a branch; the rest is selects.  metal_acc: as for absorb_material.
__device__ __forceinline__ bool absorb_one(const DevState &st, double charge, Part &p, const RowRuns rr = RowRuns{},
                                           double *metal_acc = nullptr)
{
    const int nz = st.nz, nr = st.nr;
    const bool out = p.Z <= 0.0 || p.Z >= (double)(nz - 1) || p.R >= (double)(nr - 1);
    const bool refl = !out && p.R <= 0.0;
    p.R = refl ? fabs(p.R) : p.R;
    p.Vr = refl ? fabs(p.Vr) : p.Vr;
    p.Ar = refl ? fabs(p.Ar) : p.Ar;
    // inside the domain the clamps only matter for a non-finite position,
    // which must not read outside the geometry arrays
    const int NzL = min(max((int)ceil(p.Z), 1), nz - 1);
    const int NrL = min(max((int)ceil(p.R), 1), nr - 1);
    const int jr = NrL - 1 - rr.r0;
    // run of row r0 + jr, picked by the bits of jr (two levels of selects)
    const bool odd = (jr & 1) != 0, high = (jr & 2) != 0;
    const int lo = high ? (odd ? rr.v3.x : rr.v2.x) : (odd ? rr.v1.x : rr.v0.x);
    const int hi = high ? (odd ? rr.v3.y : rr.v2.y) : (odd ? rr.v1.y : rr.v0.y);
    const bool vac = jr >= 0 && jr < rr.n && NzL - 1 >= lo && NzL - 1 < hi;
    bool del = out;
    if (!out && !refl && !vac) del = absorb_material(st, charge, p, NzL, NrL, metal_acc);
    return del;
}

// ---------------------------------------------------------------- MCC
struct Mc {
    double miu, mass_ratio, energy;
    double gx, gy, gz, gper, g;
    Part poi, pot;
};

// VecRanXYZInit, R/src/pic/particle/ParticleOne.f90:108-123
__device__ __forceinline__ void vec_ran_xyz(Part &p, double V, Rng &rg)
{
    double R = rg.next();
    double CosTheta = 1.0 - 2.0 * R;
    double SinTheta = sqrt(1.0 - CosTheta * CosTheta);
    R = rg.next();
    double Phi = 2.0 * kPI * R;
    double s, c;
    sincos(Phi, &s, &c);
    p.Vz = V * CosTheta;
    p.Vr = V * SinTheta * c;
    p.Vt = V * SinTheta * s;
}

// VelMaxInit, R/src/pic/particle/ParticleOne.f90:76-95
// (Beta = 1/(kB*Temperature) comes from the caller's table.)
__device__ __forceinline__ void vel_max_init(Part &p, double Mass, double Beta, Rng &rg)
{
    double FuncA = 1.0, FuncB = 0.0, R = 0.5, lr = 0.0;
    while (FuncA > FuncB) {
        R = rg.next();
        FuncA = R * R;
        R = rg.next();
        lr = log(R);
        FuncB = -kEXP * R * lr;
    }
    double V = sqrt(-3.0 * lr / Beta / Mass);
    vec_ran_xyz(p, V, rg);
}

// Updater, R/src/mc/TypeMCC.f90:60-92
__device__ __forceinline__ void mc_update(Mc &m, double sm, const DevTable &t, Rng &rg)
{
    const double gm = t.gas_mass;
    m.pot = m.poi;
    m.miu = t.miu;                 // sm * gm / (sm + gm)
    m.mass_ratio = t.mass_ratio;   // gm / (sm + gm)
    vel_max_init(m.pot, gm, t.beta, rg);
    m.gx = m.poi.Vr - m.pot.Vr;
    m.gy = m.poi.Vt - m.pot.Vt;
    m.gz = m.poi.Vz - m.pot.Vz;
    double Gtemp = m.gy * m.gy + m.gz * m.gz;
    m.gper = sqrt(Gtemp);
    double G2 = Gtemp + m.gx * m.gx;
    m.g = sqrt(G2);
    m.energy = 0.5 * m.miu * G2 / kJTOEV;
    m.pot.Ar = m.pot.Ar * sm / gm;
    m.pot.Az = m.pot.Az * sm / gm;
}

// VelocityUpdater, R/src/mc/TypeMCC.f90:94-112
__device__ __forceinline__ void mc_vel_update(Mc &m, double VF)
{
    m.gx = VF * m.gx; m.gy = VF * m.gy; m.gz = VF * m.gz;
    m.gper = VF * m.gper; m.g = VF * m.g;
    m.poi.Vr = VF * m.poi.Vr; m.poi.Vt = VF * m.poi.Vt; m.poi.Vz = VF * m.poi.Vz;
}

// PostCollisionVelocity, R/src/mc/TypeMCC.f90:114-146 with IsotropicCosKai or
// CosKaiVahedi (R/src/mc/MCCChi.f90:5-25)
__device__ __forceinline__ void post_collision(Mc &m, bool vahedi, Rng &rg)
{
    if (fabs(m.gper) < kMINREAL) {
        vec_ran_xyz(m.poi, m.g, rg);
        return;
    }
    double CosTheta;
    if (vahedi) {
        if (m.energy < kMINREAL) {
            CosTheta = 1.0;
        } else {
            double R = rg.next();
            CosTheta = (m.energy + 2.0 - 2.0 * pow(m.energy + 1.0, R)) / m.energy;
        }
    } else {
        double R = rg.next();
        CosTheta = 1.0 - 2.0 * R;
    }
    double FcosTheta = 1.0 - CosTheta;
    double SinTheta = sqrt(1.0 - CosTheta * CosTheta);
    double R = rg.next();
    double APhi = 2.0 * kPI * R;
    double SinPhi, CosPhi;
    sincos(APhi, &SinPhi, &CosPhi);
    const double gx = m.gx, gy = m.gy, gz = m.gz, g = m.g, gper = m.gper;
    double hx = gper * CosPhi;
    double hy = -(gx * gy * CosPhi + g * gz * SinPhi) / gper;
    double hz = -(gx * gz * CosPhi - g * gy * SinPhi) / gper;
    m.poi.Vr = m.poi.Vr - m.mass_ratio * (gx * FcosTheta + hx * SinTheta);
    m.poi.Vt = m.poi.Vt - m.mass_ratio * (gy * FcosTheta + hy * SinTheta);
    m.poi.Vz = m.poi.Vz - m.mass_ratio * (gz * FcosTheta + hz * SinTheta);
}

// SelectProbility, R/src/mc/TypeMCC.f90:173-216.  R is the uniform of the
// reaction choice, which the caller draws before Updater's target sampling
// (draw 1 of the particle's stream for this gas): the particle pass can then
// screen out null events before they are queued.
__device__ __forceinline__ int select_probability(const Mc &m, const DevTable &t, double R)
{
    const int Nr = t.nreaction, Ns = t.nsigma;
    const double E = m.energy;
    const auto *__restrict__ P = gp(t.prob);
    const auto *__restrict__ TH = gp(t.threshold);
    double lx = (log10(E) - t.emin) / t.einterval;
    int N = (E > 0.0) ? (int)ceil(lx) : 0;
    int mode;
    double S1 = 0.0, S2 = 0.0;
    decltype(P) a, b = nullptr;
    if (N < 1) {
        mode = 0; a = P;
    } else if (N < Ns) {
        mode = 1;
        S1 = (double)N - (log10(E) - t.emin) / t.einterval;
        S2 = 1.0 - S1;
        a = P + (size_t)(N - 1) * Nr;
        b = P + (size_t)N * Nr;
    } else {
        mode = 2; a = P + (size_t)(Ns - 1) * Nr;
    }
    auto value = [&](int j) {
        double v;
        if (mode == 0) v = a[j];
        else if (mode == 1) v = a[j] * S1 + b[j] * S2;
        else v = a[j] * E / t.emax;
        if (E < TH[j]) v = 0.0;
        return v;
    };
    // A cumulative table (columns nondecreasing in j, as ProbilityNonReactive
    // builds them) interpolates to values nondecreasing in j (the weights
    // are nonnegative and rounding is monotone), so R >= the last value
    // misses every reaction: the null event is known from one lookup.
    if (t.cumulative && !(E < TH[Nr - 1]) && !(R < value(Nr - 1))) return 0;
    for (int j = 0; j < Nr; ++j)
        if (R < value(j)) return j + 1;
    return 0;
}

// SelectProbilityReactive, R/src/mc/TypeMCC.f90:218-236; R (the draw of
// beta) is drawn before Updater, as SelectProbility's is, and the
// non-reactive choice takes the stream's next draw.
__device__ __forceinline__ int select_probability_reactive(const Mc &m, const DevTable &t, double R, Rng &rg)
{
    double beta = t.beta_max * sqrt(R);
    if (beta <= 1.0) {
        if (m.energy <= gp(t.threshold)[1]) return 1;
        return select_probability(m, t, rg.next());
    }
    return 1;
}

// (log10(E) - emin) / einterval within t.nb_eps bins, for E > 0 normal:
// log2 E = exponent + log2 of the mantissa in [0.5, 1), the latter from the
// hardware single-precision log2 (error below 1e-6 there, which nb_eps
// covers together with the rounding of the float conversion).
__device__ __forceinline__ double approx_lx(const DevTable &t, double e)
{
    int ex;
    const double mt = frexp(e, &ex);
    const double l2 = (double)ex + (double)__log2f((float)mt);
    return (l2 * 0.30102999566398120 - t.emin) * t.nb_inv_einterval;
}

// True only when SelectProbility (models 1 and 2) returns 0 for every target
// velocity VelMaxInit can draw: R is at or above every value the table can
// interpolate over the relative energies the particle can reach.  vz, vr, vt
// are the particle's velocity after VelRes(VFactor) (m/s).  The relative
// speed lies within nb_vmax of the particle's speed, so the energy within
// [e_lo, e_hi]; the bins of that range come from approx_lx widened by
// nb_eps, and the sparse table gives their bound with two loads.  The
// threshold test of SelectProbility only lowers values, so it is left out.
__device__ __forceinline__ bool surely_null(const DevTable &t, double vz, double vr, double vt, double R)
{
    const double v2 = vz * vz + vr * vr + vt * vt;
    if (!(v2 < 1e300)) return false;
    const double v = sqrt(v2);
    const double glo = v - t.nb_vmax, ghi = v + t.nb_vmax;
    const double ehi = t.nb_ecoef * ghi * ghi * (1.0 + 1e-12);
    const double elo = glo > 0.0 ? t.nb_ecoef * glo * glo * (1.0 - 1e-12) : 0.0;
    const int Ns = t.nsigma;
    int lo = 0, hi = Ns;
    if (elo > 1e-300) lo = (int)fmin(fmax(ceil(approx_lx(t, elo) - t.nb_eps), 0.0), (double)Ns);
    if (ehi < 1e300) hi = (int)fmin(fmax(ceil(approx_lx(t, ehi) + t.nb_eps), 0.0), (double)Ns);
    if (hi < lo) return false;
    const int L = 31 - __clz(hi - lo + 1);
    const auto *row = gp(t.nb_tab) + (size_t)L * (size_t)(Ns + 1);
    double bound = fmax(row[lo], row[hi - (1 << L) + 1]);
    if (hi == Ns) bound = bound * fmax(1.0, ehi * t.nb_inv_emax) * (1.0 + 1e-12);
    return R >= bound;
}

// Cell (ceiling(Z), ceiling(R)) of a particle, clamped into the grid's
// cells, as a 0-based index (iz - 1) + (ir - 1) * (nz - 1).
__device__ __forceinline__ unsigned long long cell_of(const DevState &st, double z, double r)
{
    int iz = min(max((int)ceil(z), 1), st.nz - 1);
    int ir = min(max((int)ceil(r), 1), st.nr - 1);
    return (unsigned long long)(iz - 1) + (unsigned long long)(ir - 1) * (st.nz - 1);
}

}  // namespace ipm
