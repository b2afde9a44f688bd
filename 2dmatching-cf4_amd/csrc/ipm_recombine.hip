// Volume recombination between particle species: executingRecombination2D
// and deletingParticles, R/src/mc/Recombination.F90:164-273.
//
// The reference's deletion is sequential inside each cell: the particles of
// a bundle are visited from the last storage slot down, each cell's budget
// of physical particles to remove shrinking by every deleted weight, and the
// first particle heavier than what is left is deleted with probability
// budget / WQ.  Which particles go therefore depends on the visiting order
// within a cell, and on nothing else.  Here every involved bundle is first
// reordered by a stable sort on the cell (radix sort of (cell, slot) pairs,
// dead slots last), which keeps the slot order inside each cell; one thread
// per cell then walks its contiguous segment from the top, exactly the
// reference's loop restricted to that cell.  Cell sums run in ascending slot
// order as the reference's accumulation loop does, so densities, budgets and
// decisions are bit for bit those of the oracle given the same storage.
// Deleted particles are marked dead in place (the reference's DelOne
// swaps the last particle into the hole; a set of survivors does not depend
// on that, but later passes of the same call visit a bundle in the order it
// had when the call started).
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cmath>
#include "ipm_ctx.hpp"

namespace {

constexpr int kBlock = 256;

constexpr uint32_t kRecombTag = 0x52430000u;   // apart from the collision and merge tags

__device__ __forceinline__ bool rc_dead(double z) { return z < -1.0e299; }

struct Ptr8 {
    double *p[8];
};

// Cell of every slot of species s (ncell for a dead slot) and the slot index.
__global__ void rc_keys_kernel(const DevState *__restrict__ stp, int s, long long n, unsigned ncell,
                               unsigned *__restrict__ keys, unsigned *__restrict__ vals)
{
    const DevState &st = *stp;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double z = st.f[s][0][i];
    keys[i] = rc_dead(z) ? ncell : (unsigned)cell_of(st, z, st.f[s][1][i]);
    vals[i] = (unsigned)i;
}

// dst[k] = src[vals[k]] for the nlive live particles, in sorted order.
__global__ void rc_gather_kernel(const DevState *__restrict__ stp, int s, long long nlive,
                                 const unsigned *__restrict__ vals, Ptr8 dst)
{
    const DevState &st = *stp;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlive) return;
    const unsigned j = vals[k];
#pragma unroll
    for (int f = 0; f < 8; ++f) dst.p[f][k] = st.f[s][f][j];
}

__global__ void rc_set_count_kernel(DevCounters *ctr, int s, long long nlive)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctr->n[s] = (unsigned long long)nlive;
        ctr->n_new[s] = 0;
        ctr->n_dead[s] = 0;
    }
}

// Segment [start[c], end[c]) of every occupied cell c in the sorted keys
// (n of them, the nlive live ones first).  A live count that disagrees with
// the keys raises overflow bit 8.
__global__ void rc_segments_kernel(const unsigned *__restrict__ keys, long long n, long long nlive, unsigned ncell,
                                   unsigned *__restrict__ start, unsigned *__restrict__ end, DevCounters *ctr)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlive) return;
    const unsigned key = keys[k];
    if (k == nlive - 1 && (key >= ncell || (nlive < n && keys[nlive] != ncell))) {
        atomicOr(&ctr->overflow, 8);
        return;
    }
    if (key >= ncell) return;
    if (k == 0 || keys[k - 1] != key) start[key] = (unsigned)k;
    if (k == nlive - 1 || keys[k + 1] != key) end[key] = (unsigned)(k + 1);
}

// 1/cell volume and cell volume of radial cell ir, as initRecombination2D
// forms them (Recombination.F90:133-137): PI * dr * dr * dz * (i*i - (i-1)*(i-1)).
__device__ __forceinline__ double rc_volume(double dx, int ir)
{
    return kPI * dx * dx * dx * (double)(ir * ir - (ir - 1) * (ir - 1));
}

// Number density of species s in each cell and, when want_t, the inverse of
// the mean energy in eV (Recombination.F90:175-206).
__global__ void rc_density_kernel(const DevState *__restrict__ stp, int s, int want_t,
                                  const unsigned *__restrict__ start, const unsigned *__restrict__ end,
                                  double mass, double *__restrict__ dens, double *__restrict__ inv_t)
{
    const DevState &st = *stp;
    const int ncz = st.nz - 1;
    const long long ncell = (long long)ncz * (st.nr - 1);
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int ir = (int)(c / ncz) + 1;
    double *const *F = st.f[s];
    double sw = 0.0, se = 0.0;
    const unsigned k0 = start[c], k1 = end[c];
    for (unsigned k = k0; k < k1; ++k) {
        const double w = F[7][k];
        sw = sw + w;
        if (want_t) {
            const double vz = F[2][k], vr = F[3][k], vt = F[4][k];
            const double energy = vz * vz + vr * vr + vt * vt;
            se = se + energy * w;
        }
    }
    const double inv_vol = 1.0 / rc_volume(st.dx, ir);
    const double d = sw * inv_vol;
    dens[c] = d;
    if (want_t) {
        const double t = se * inv_vol * 0.5 * mass * st.vfactor * st.vfactor;
        inv_t[c] = (d > 0.0 && t > 0.0) ? 1.0 / (t / kJTOEV / d) : 0.0;
    }
}

// deletingParticles (Recombination.F90:246-273) of species x in every cell,
// with the cell's budget recb_neg = recb_pos of reaction type `type`
// (Recombination.F90:213-235).  Weights are positive, so once the budget is
// spent no later particle of the cell can be deleted and the walk stops.
__global__ void rc_delete_kernel(const DevState *__restrict__ stp, DevCounters *__restrict__ ctr, int x, int type,
                                 double rate, double time_step, const double *__restrict__ dn,
                                 const double *__restrict__ dp, const double *__restrict__ tn,
                                 const double *__restrict__ tp, const unsigned *__restrict__ start,
                                 const unsigned *__restrict__ end, uint32_t tag,
                                 unsigned long long *__restrict__ ndel)
{
    const DevState &st = *stp;
    const int ncz = st.nz - 1;
    const long long ncell = (long long)ncz * (st.nr - 1);
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const unsigned k0 = start[c], k1 = end[c];
    if (k1 <= k0) return;
    const int ir = (int)(c / ncz) + 1;
    const double vol = rc_volume(st.dx, ir);
    double b;
    if (type == 0)
        b = dn[c] * dp[c] * rate * time_step * vol;
    else
        b = dn[c] * dp[c] * time_step * vol * rate * sqrt(tn[c]) * tp[c];
    if (!(b > 0.0)) return;
    double *const *F = st.f[x];
    const unsigned long long seed = ctr->seed;
    const uint32_t step = ctr->step;
    unsigned long long cnt = 0;
    for (long long k = (long long)k1 - 1; k >= (long long)k0; --k) {
        const double z = F[0][k];
        if (rc_dead(z)) continue;
        const double w = F[7][k];
        if (b > w) {
            b = b - w;
            F[0][k] = kDead;
            ++cnt;
        } else {
            const StateKey sk = state_key(z, F[1][k], F[2][k], F[3][k], F[4][k]);
            Rng g = make_rng(sk, seed, step, tag);
            if (g.next() < b / w) {
                b = 0.0;
                F[0][k] = kDead;
                ++cnt;
            }
        }
        if (!(b > 0.0)) break;
    }
    if (cnt) {
        atomicAdd(&ctr->n_dead[x], cnt);
        atomicAdd(&ndel[x], cnt);
    }
}

}  // namespace

struct ipm_recomb {
    unsigned *keys[2] = {}, *vals[2] = {};
    long long kcap = 0;
    void *temp = nullptr;
    size_t temp_bytes = 0;
    unsigned *seg = nullptr;               // start, end: 2 * ncell per species
    double *dens = nullptr, *inv_t = nullptr;   // ncell per species
    unsigned long long *ndel = nullptr;    // IPM_NS_MAX
};

int recomb_release(ipm_ctx *c)
{
    if (!c || !c->recomb) return IPM_OK;
    ipm_recomb *r = c->recomb;
    void *ptrs[] = {r->keys[0], r->keys[1], r->vals[0], r->vals[1], r->temp, r->seg, r->dens, r->inv_t, r->ndel};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete r;
    c->recomb = nullptr;
    return IPM_OK;
}

static int recomb_buffers(ipm_ctx *c, long long nmax, int bits)
{
    if (!c->recomb) c->recomb = new ipm_recomb();
    ipm_recomb *r = c->recomb;
    const size_t ncell = (size_t)(c->nz - 1) * (c->nr - 1);
    if (!r->seg) {
        HIPCHK(hipMalloc(&r->seg, sizeof(unsigned) * 2 * ncell * IPM_NS_MAX));
        HIPCHK(hipMalloc(&r->dens, sizeof(double) * ncell * IPM_NS_MAX));
        HIPCHK(hipMalloc(&r->inv_t, sizeof(double) * ncell * IPM_NS_MAX));
        HIPCHK(hipMalloc(&r->ndel, sizeof(unsigned long long) * IPM_NS_MAX));
    }
    if (nmax > r->kcap) {
        for (int b = 0; b < 2; ++b) {
            if (r->keys[b]) HIPCHK(hipFree(r->keys[b]));
            if (r->vals[b]) HIPCHK(hipFree(r->vals[b]));
            r->keys[b] = r->vals[b] = nullptr;
        }
        r->kcap = 0;
        for (int b = 0; b < 2; ++b) {
            HIPCHK(hipMalloc(&r->keys[b], sizeof(unsigned) * (size_t)nmax));
            HIPCHK(hipMalloc(&r->vals[b], sizeof(unsigned) * (size_t)nmax));
        }
        r->kcap = nmax;
    }
    size_t need = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, r->keys[0], r->keys[1], r->vals[0], r->vals[1],
                                              (int)nmax, 0, bits, c->stream));
    if (need > r->temp_bytes) {
        if (r->temp) HIPCHK(hipFree(r->temp));
        r->temp = nullptr;
        r->temp_bytes = 0;
        HIPCHK(hipMalloc(&r->temp, need));
        r->temp_bytes = need;
    }
    return IPM_OK;
}

int ipm_recombine(ipm_ctx *c, int nreaction, const int *reaction_list, const double *rate, double time_step,
                  int64_t *deleted)
{
    int rc = check_ready(c, false, false);
    if (rc) return rc;
    if (nreaction < 0 || nreaction > 4095) return fail(IPM_EINVAL, "invalid recombination reaction count %d", nreaction);
    if (nreaction > 0 && (!reaction_list || !rate)) return fail(IPM_EINVAL, "null reaction list or rates");
    if (!std::isfinite(time_step)) return fail(IPM_EINVAL, "non-finite recombination time step");
    bool involved[IPM_NS_MAX] = {}, want_t[IPM_NS_MAX] = {};
    for (int i = 0; i < nreaction; ++i) {
        const int neg = reaction_list[i], pos = reaction_list[nreaction + i], type = reaction_list[2 * nreaction + i];
        if (neg < 0 || neg >= c->ns || pos < 0 || pos >= c->ns)
            return fail(IPM_EINVAL, "The specified particle index is invalid in recombination reactions (reaction %d).",
                        i + 1);
        if (type != 0 && type != 1) return fail(IPM_EINVAL, "invalid recombination type %d (reaction %d)", type, i + 1);
        if (!std::isfinite(rate[i])) return fail(IPM_EINVAL, "non-finite recombination rate (reaction %d)", i + 1);
        involved[neg] = involved[pos] = true;
        if (type == 1) want_t[neg] = want_t[pos] = true;
    }
    DevCounters ctr;
    if ((rc = read_counters(c, &ctr))) return rc;
    const unsigned ncell = (unsigned)((c->nz - 1) * (c->nr - 1));
    int bits = 1;
    while ((1ull << bits) <= (unsigned long long)ncell) ++bits;
    long long nmax = 1;
    for (int s = 0; s < c->ns; ++s)
        if (involved[s]) {
            if (c->h.cap[s] >= 0xFFFFFFFFll) return fail(IPM_EINVAL, "species %d capacity too large for recombination", s);
            nmax = std::max<long long>(nmax, (long long)ctr.n[s]);
        }
    if ((rc = recomb_buffers(c, nmax, bits))) return rc;
    ipm_recomb *r = c->recomb;
    const unsigned cblk = (ncell + kBlock - 1) / kBlock;
    HIPCHK(hipMemsetAsync(r->ndel, 0, sizeof(unsigned long long) * IPM_NS_MAX, c->stream));
    for (int s = 0; s < c->ns; ++s) {
        if (!involved[s]) continue;
        const long long n = (long long)ctr.n[s], nlive = n - (long long)ctr.n_dead[s];
        unsigned *start = r->seg + (size_t)2 * ncell * s, *end = start + ncell;
        HIPCHK(hipMemsetAsync(start, 0, sizeof(unsigned) * 2 * ncell, c->stream));
        if (n > 0) {
            const unsigned pblk = (unsigned)((n + kBlock - 1) / kBlock);
            hipLaunchKernelGGL(rc_keys_kernel, dim3(pblk), dim3(kBlock), 0, c->stream, c->d_state, s, n, ncell,
                               r->keys[0], r->vals[0]);
            HIPCHK(hipGetLastError());
            size_t tb = r->temp_bytes;
            HIPCHK(hipcub::DeviceRadixSort::SortPairs(r->temp, tb, r->keys[0], r->keys[1], r->vals[0], r->vals[1],
                                                      (int)n, 0, bits, c->stream));
            const int other = 1 - c->cur[s];
            Ptr8 dst;
            for (int f = 0; f < 8; ++f) dst.p[f] = c->h.alt[s][f];
            if (nlive > 0) {
                const unsigned gblk = (unsigned)((nlive + kBlock - 1) / kBlock);
                hipLaunchKernelGGL(rc_gather_kernel, dim3(gblk), dim3(kBlock), 0, c->stream, c->d_state, s, nlive,
                                   r->vals[1], dst);
                hipLaunchKernelGGL(rc_segments_kernel, dim3(gblk), dim3(kBlock), 0, c->stream, r->keys[1], n, nlive,
                                   ncell, start, end, c->d_ctr);
            }
            hipLaunchKernelGGL(rc_set_count_kernel, dim3(1), dim3(64), 0, c->stream, c->d_ctr, s, nlive);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(c->stream));
            set_buffers(c, s, other);
            if ((rc = sync_state(c))) return rc;
        }
        hipLaunchKernelGGL(rc_density_kernel, dim3(cblk), dim3(kBlock), 0, c->stream, c->d_state, s, want_t[s] ? 1 : 0,
                           start, end, c->h.mass[s], r->dens + (size_t)ncell * s, r->inv_t + (size_t)ncell * s);
        HIPCHK(hipGetLastError());
    }
    for (int i = 0; i < nreaction; ++i) {
        const int neg = reaction_list[i], pos = reaction_list[nreaction + i], type = reaction_list[2 * nreaction + i];
        const double *dn = r->dens + (size_t)ncell * neg, *dp = r->dens + (size_t)ncell * pos;
        const double *tn = r->inv_t + (size_t)ncell * neg, *tp = r->inv_t + (size_t)ncell * pos;
        for (int role = 0; role < 2; ++role) {
            const int x = role == 0 ? neg : pos;
            const unsigned *start = r->seg + (size_t)2 * ncell * x, *end = start + ncell;
            const uint32_t tag = kRecombTag | ((uint32_t)i << 4) | ((uint32_t)role << 3) | (uint32_t)x;
            hipLaunchKernelGGL(rc_delete_kernel, dim3(cblk), dim3(kBlock), 0, c->stream, c->d_state, c->d_ctr, x, type,
                               rate[i], time_step, dn, dp, tn, tp, start, end, tag, r->ndel);
            HIPCHK(hipGetLastError());
        }
    }
    unsigned long long nd[IPM_NS_MAX];
    HIPCHK(hipMemcpyAsync(nd, r->ndel, sizeof nd, hipMemcpyDeviceToHost, c->stream));
    if ((rc = read_counters(c, &ctr))) return rc;
    if (deleted)
        for (int s = 0; s < c->ns; ++s) deleted[s] = (int64_t)nd[s];
    return IPM_OK;
}
