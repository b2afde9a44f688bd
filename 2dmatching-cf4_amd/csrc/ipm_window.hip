// Differentiable collisionless window over the particle push: the forward
// (record or replay of absorption events) and its hand-written adjoint.
// The reference has no adjoint; the forward step is MoveESIm
// (R/src/pic/particle/ParticleBundle.f90:685-745) followed by the absorption
// rule of MaterialAborptionParticle (R/src/pic/particleboundary/ParticleMaterial.F90:22-85)
// and a P2CES-weighted loss deposit (R/src/pic/field/FieldOne.f90:151-166).
//
// Field at step n: E = E0 + c[n] * E1 (two Ez/Er grids, stored as
// interleaved pairs).  Loss of step n, on the state after the push and the
// absorption of step n:
//     l[n] = sum_i alpha[s_i] * WQ_i * Vz_i * M(Z_i, R_i),
//     M(Z, R) = sum_nodes mu(node) * W_node(Z, R)
// with the bilinear node weights W of P2CES.  With mu = mask * point_inv_volume
// and alpha[s] = dt * w_s * Charge_s * VFactor this is dt * sum_s w_s *
// sum_nodes mask * JzOne_s, JzOne_s being WeightingParticleField's current
// density with the particle's own velocity.
//
// The field does not depend on the particles inside the window, so every
// particle's trajectory is independent: one launch advances every particle
// through one checkpoint segment of K steps with its state in registers, and
// writes the state at the segment end as the next checkpoint.  The adjoint
// of segment j recomputes the segment from checkpoint j, keeping each step's
// pre-push state on a tape, then runs the reverse sweep with the adjoint
// state in registers.  Absorption events are frozen: a recording forward
// stores each particle's absorption step, and replays (and the adjoint)
// delete the particle at that step.
#include <string.h>
#include <algorithm>
#include <cmath>
#include <vector>
#include "ipm_ctx.hpp"

namespace {

constexpr int kWBlock = 256;
constexpr int kKMax = 64;     // largest checkpoint interval (per-step LDS sums)
constexpr int kNS = 7;        // Z R Vz Vr Vt Az Ar (WQ is constant)
constexpr int kTape = 5;      // Z R Vr Vt Ar: what the adjoint of a push reads

struct WinArgs {
    int nz, nr, ns, T, K, nseg;
    long long N;
    long long off[IPM_NS_MAX + 1];
    double efactor[IPM_NS_MAX];
    double alpha[IPM_NS_MAX];
    double zlo, zhi, rlo, rhi;     // M and its gradient vanish unless zlo < Z <= zhi, rlo < R <= rhi
    const Pair2 *Ezp, *Erp;
    const double *mu;
    const int *cell_material, *mat_type;
    int ndiel;
    int prefilter;                 // adjoint: skip wavefronts that never reach the support
    double *ck;                    // (nseg + 1) checkpoints of kNS arrays of N
    const double *wq;              // N
    int *death;                    // N: absorption step; T = survives; -1 = empty slot
    double *tape;                  // K * kTape arrays of N
    double *lam;                   // kNS arrays of N
    const double *c;               // T
    double *acc;                   // T: loss (forward) or dl/dc (adjoint)
    const double *g;               // T: weights of the per-step losses (adjoint)
};

struct Adj {
    double Z, R, Vz, Vr, Vt, Az, Ar;
};

__device__ __forceinline__ bool dead_slot(double z) { return z < -1.0e299; }

__device__ __forceinline__ int species_of(const WinArgs &a, long long i)
{
    int s = 0;
    for (int q = 1; q < a.ns; ++q)
        if (i >= a.off[q]) s = q;
    return s;
}

__device__ __forceinline__ void load_state(const double *base, long long N, long long i, Part &p)
{
    p.Z = gp(base)[i]; p.R = gp(base + N)[i]; p.Vz = gp(base + 2 * N)[i]; p.Vr = gp(base + 3 * N)[i];
    p.Vt = gp(base + 4 * N)[i]; p.Az = gp(base + 5 * N)[i]; p.Ar = gp(base + 6 * N)[i];
}
__device__ __forceinline__ void store_state(double *base, long long N, long long i, const Part &p)
{
    gp(base)[i] = p.Z; gp(base + N)[i] = p.R; gp(base + 2 * N)[i] = p.Vz; gp(base + 3 * N)[i] = p.Vr;
    gp(base + 4 * N)[i] = p.Vt; gp(base + 5 * N)[i] = p.Az; gp(base + 6 * N)[i] = p.Ar;
}

// Absorption rule without tallies, R/src/pic/particleboundary/ParticleMaterial.F90:22-41:
// true if the particle is deleted; reflection at R <= 0 edits R, Vr, Ar.
__device__ __forceinline__ bool absorbed(const WinArgs &a, Part &p)
{
    const int nz = a.nz, nr = a.nr;
    if (p.Z <= 0.0 || p.Z >= (double)(nz - 1) || p.R >= (double)(nr - 1)) return true;
    if (p.R <= 0.0) {
        p.R = fabs(p.R); p.Vr = fabs(p.Vr); p.Ar = fabs(p.Ar);
        return false;
    }
    const int NzL = min(max((int)ceil(p.Z), 1), nz - 1);
    const int NrL = min(max((int)ceil(p.R), 1), nr - 1);
    const int m = gp(a.cell_material)[(size_t)(NzL - 1) + (size_t)(NrL - 1) * (nz - 1)];
    const int t = gp(a.mat_type)[m + a.ndiel];
    return t == kMatMetal || t == kMatDiel;
}

__device__ __forceinline__ bool in_support(const WinArgs &a, const Part &p)
{
    return p.Z > a.zlo && p.Z <= a.zhi && p.R > a.rlo && p.R <= a.rhi;
}

// M(Z, R) and its gradient, with the node weights and index guards of the
// deposit: R == 0 exactly puts the RU weights on row 1 (the RL row has weight 0).
__device__ __forceinline__ double mask_at(const WinArgs &a, double Z, double R, double &dz, double &dr)
{
    int NzL = (int)ceil(Z);
    const double ZL = (double)NzL - Z, ZU = 1.0 - ZL;
    int NrL = (int)ceil(R);
    const double RL = (double)NrL - R, RU = 1.0 - RL;
    NzL = min(max(NzL, 1), a.nz - 1);
    const bool row0 = NrL < 1;
    NrL = min(max(NrL, 1), a.nr - 1);
    const double *m = a.mu + (size_t)(NzL - 1) + (size_t)(NrL - 1) * a.nz;
    const double mLL = gp(m)[0], mUL = gp(m)[1];
    double mLU = gp(m)[a.nz], mUU = gp(m)[a.nz + 1];
    if (row0) { mLU = mLL; mUU = mUL; }
    dz = (mUL * RL + mUU * RU) - (mLL * RL + mLU * RU);
    dr = (mLU * ZL + mUU * ZU) - (mLL * ZL + mUL * ZU);
    return mLL * ZL * RL + mLU * ZL * RU + mUL * ZU * RL + mUU * ZU * RU;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ void lds_add_d(double *a, double v)
{
    __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Adjoint of push_core with the PairField gather: on entry l holds the
// adjoint of the pushed state, on exit that of the state q before the push;
// returns the derivative with respect to the field coefficient c and, in
// E2out, the scaled Ez of the push.  Only q.Z, q.R, q.Vr, q.Vt, q.Ar are
// read: the push is affine in Vz and Az.
__device__ __forceinline__ double push_adjoint(const WinArgs &a, double F, double c, const Part &q, Adj &l,
                                               double &E2out)
{
    const Gather g = gather_at(a.nz, a.nr, q.Z, q.R);
    const Pair2 a00 = ld_pair(a.Erp, g.kr), a10 = ld_pair(a.Erp, g.kr + 1);
    const Pair2 a01 = ld_pair(a.Erp, g.kr + g.ldr), a11 = ld_pair(a.Erp, g.kr + g.ldr + 1);
    const Pair2 b00 = ld_pair(a.Ezp, g.kz0), b10 = ld_pair(a.Ezp, g.kz0 + 1);
    const Pair2 b01 = ld_pair(a.Ezp, g.kz1), b11 = ld_pair(a.Ezp, g.kz1 + 1);
    const double A00 = a00.x + c * a00.y, A10 = a10.x + c * a10.y, A01 = a01.x + c * a01.y, A11 = a11.x + c * a11.y;
    const double B00 = b00.x + c * b00.y, B10 = b10.x + c * b10.y, B01 = b01.x + c * b01.y, B11 = b11.x + c * b11.y;
    const double Z1 = g.Z1, Z2 = g.Z2, Z1H = g.Z1H, Z2H = g.Z2H, R1 = g.R1, R2 = g.R2, R1H = g.R1H, R2H = g.R2H;

    const double E1 = (A00 * Z1 * R1H + A10 * Z2 * R1H + A01 * Z1 * R2H + A11 * Z2 * R2H) * F;
    E2out = (B00 * Z1H * R1 + B10 * Z2H * R1 + B01 * Z1H * R2 + B11 * Z2H * R2) * F;
    // d/dZ of Z1, Z1H is -1 and of Z2, Z2H +1; likewise in R
    const double dE1dz = ((A10 - A00) * R1H + (A11 - A01) * R2H) * F;
    const double dE1dr = ((A01 * Z1 + A11 * Z2) - (A00 * Z1 + A10 * Z2)) * F;
    const double dE2dz = ((B10 - B00) * R1 + (B11 - B01) * R2) * F;
    const double dE2dr = ((B01 * Z1H + B11 * Z2H) - (B00 * Z1H + B10 * Z2H)) * F;
    const double dE1dc = (a00.y * Z1 * R1H + a10.y * Z2 * R1H + a01.y * Z1 * R2H + a11.y * Z2 * R2H) * F;
    const double dE2dc = (b00.y * Z1H * R1 + b10.y * Z2H * R1 + b01.y * Z1H * R2 + b11.y * Z2H * R2) * F;

    // forward intermediates of the r-theta rotation
    const double vr = q.Vr + E1;
    const double rr = q.R + E1;
    const double ar = (q.Ar + E1) * 0.5;
    const double Vx = vr + ar;
    const double Vy = q.Vt;
    const double X = rr + Vx;
    const double Y = Vy;
    const double P = sqrt(X * X + Y * Y);

    double gX, gY, gVx, gVy, gar;
    if (P < kMINREAL) {
        // axis branch: Vr' = |(Vx, Vy)|, Vt' = MinReal, Ar' = |ar|, R' = P
        const double h = sqrt(Vx * Vx + Vy * Vy);
        gVx = h > 0.0 ? l.Vr * Vx / h : 0.0;
        gVy = h > 0.0 ? l.Vr * Vy / h : 0.0;
        gar = ar >= 0.0 ? l.Ar : -l.Ar;
        gX = P > 0.0 ? l.R * X / P : 0.0;
        gY = P > 0.0 ? l.R * Y / P : 0.0;
    } else {
        const double C = X / P, S = Y / P;
        const double gC = l.Vr * Vx + l.Vt * Vy + l.Ar * ar;
        const double gS = l.Vr * Vy - l.Vt * Vx;
        gX = (S / P) * (gC * S - gS * C) + l.R * C;
        gY = (C / P) * (gS * C - gC * S) + l.R * S;
        gVx = l.Vr * C - l.Vt * S;
        gVy = l.Vr * S + l.Vt * C;
        gar = l.Ar * C;
    }
    gVy += gY;          // Y = Vy
    gVx += gX;          // X = rr + Vx
    const double grr = gX;
    const double gvr = gVx;   // Vx = vr + ar
    gar += gVx;
    // z: zz' = (Z + E2) + vz', vz' = (Vz + E2) + az, az = (Az + E2) / 2
    const double gvz2 = l.Vz + l.Z;
    const double gaz = l.Az + gvz2;
    const double gE2 = 0.5 * gaz + gvz2 + l.Z;
    // r: rr = R + E1, vr = Vr + E1, ar = (Ar + E1) / 2
    const double gE1 = grr + gvr + 0.5 * gar;

    Adj o;
    o.Z = l.Z + gE1 * dE1dz + gE2 * dE2dz;
    o.R = grr + gE1 * dE1dr + gE2 * dE2dr;
    o.Vz = gvz2;
    o.Vr = gvr;
    o.Vt = gVy;
    o.Az = 0.5 * gaz;
    o.Ar = 0.5 * gar;
    l = o;
    return gE1 * dE1dc + gE2 * dE2dc;
}

// One checkpoint segment of the forward for every particle.  Reads
// checkpoint j, writes checkpoint j + 1 and the per-step losses.
__global__ void __launch_bounds__(kWBlock) win_forward_kernel(WinArgs a, int j, int record)
{
    __shared__ double sl[kKMax];
    const int n0 = j * a.K, n1 = min(a.T, n0 + a.K);
    for (int t = threadIdx.x; t < kKMax; t += kWBlock) sl[t] = 0.0;
    __syncthreads();
    const long long N = a.N;
    const double *src = a.ck + (size_t)j * kNS * N;
    double *dst = a.ck + (size_t)(j + 1) * kNS * N;
    const long long stride = (long long)gridDim.x * kWBlock;
    for (long long base = (long long)blockIdx.x * kWBlock; base < N; base += stride) {
        const long long i = base + threadIdx.x;
        const bool valid = i < N;
        Part p{};
        int d = -1, s = 0;
        double wq = 0.0;
        if (valid) {
            load_state(src, N, i, p);
            s = species_of(a, i);
            wq = gp(a.wq)[i];
            if (record && j == 0) d = dead_slot(p.Z) ? -1 : a.T;
            else d = gp(a.death)[i];
        }
        const double F = a.efactor[s], al = a.alpha[s];
        for (int n = n0; n < n1; ++n) {
            const bool live = valid && d >= n;       // pushed at step n
            const PairField fs{a.Ezp, a.Erp, a.c[n]};
            double v = 0.0;
            if (live) {
                push_core(fs, a.nz, a.nr, F, p);
                bool del;
                if (record) {
                    del = absorbed(a, p);
                    if (del) d = n;
                } else {
                    del = (n == d);
                    if (!del && p.R <= 0.0) { p.R = fabs(p.R); p.Vr = fabs(p.Vr); p.Ar = fabs(p.Ar); }
                }
                if (!del && al != 0.0 && in_support(a, p)) {
                    double dz, dr;
                    v = al * wq * p.Vz * mask_at(a, p.Z, p.R, dz, dr);
                }
            }
            if (__ballot(v != 0.0)) {
                const double w = wave_sum(v);
                if ((threadIdx.x & 63) == 0) lds_add_d(&sl[n - n0], w);
            }
        }
        if (valid) {
            store_state(dst, N, i, p);
            if (record) gp(a.death)[i] = d;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n1 - n0; t += kWBlock)
        if (sl[t] != 0.0) global_add(&a.acc[n0 + t], sl[t]);
}

// Adjoint of checkpoint segment j.  lam holds, on entry, the adjoint of the
// state at the segment end (ignored for the last segment) and, on exit, that
// of the state at the segment start.  a.acc[n] += dl/dc[n].
//
// The recompute keeps Z R Vr Vt Ar of every pre-push state on the tape; the
// reverse sweep walks the tape backwards holding the post-push state of the
// step it differentiates (the next tape entry), and carries Vz and Az back
// through the affine z-update: Az[n] = 2 Az[n+1] - E2[n],
// Vz[n] = Vz[n+1] - Az[n+1] - E2[n].
__device__ __forceinline__ void tape_store(double *t, long long N, long long i, const Part &p)
{
    gp(t)[i] = p.Z; gp(t + N)[i] = p.R; gp(t + 2 * N)[i] = p.Vr; gp(t + 3 * N)[i] = p.Vt; gp(t + 4 * N)[i] = p.Ar;
}
__device__ __forceinline__ void tape_load(const double *t, long long N, long long i, Part &p)
{
    p.Z = gp(t)[i]; p.R = gp(t + N)[i]; p.Vr = gp(t + 2 * N)[i]; p.Vt = gp(t + 3 * N)[i]; p.Ar = gp(t + 4 * N)[i];
}

__global__ void __launch_bounds__(kWBlock) win_adjoint_kernel(WinArgs a, int j)
{
    __shared__ double sg[kKMax];
    const int n0 = j * a.K, n1 = min(a.T, n0 + a.K);
    for (int t = threadIdx.x; t < kKMax; t += kWBlock) sg[t] = 0.0;
    __syncthreads();
    const long long N = a.N;
    const double *src = a.ck + (size_t)j * kNS * N;
    const bool last = (j == a.nseg - 1);
    const long long stride = (long long)gridDim.x * kWBlock;
    for (long long base = (long long)blockIdx.x * kWBlock; base < N; base += stride) {
        const long long i = base + threadIdx.x;
        const bool valid = i < N;
        int d = -1, s = 0;
        double wq = 0.0;
        if (valid) {
            s = species_of(a, i);
            d = gp(a.death)[i];
            wq = gp(a.wq)[i];
        }
        const double F = a.efactor[s], al = a.alpha[s];
        // Steps n0 <= n < nl are pushed and survive; a particle deleted at
        // step d < n1 contributes nothing from step d on.
        const bool active = valid && al != 0.0 && d > n0;
        const int nl = active ? min(n1, d) : n0;
        Adj l{0, 0, 0, 0, 0, 0, 0};
        if (active && !last && d >= n1) {
            l.Z = gp(a.lam)[i]; l.R = gp(a.lam + N)[i]; l.Vz = gp(a.lam + 2 * N)[i];
            l.Vr = gp(a.lam + 3 * N)[i]; l.Vt = gp(a.lam + 4 * N)[i]; l.Az = gp(a.lam + 5 * N)[i];
            l.Ar = gp(a.lam + 6 * N)[i];
        }
        Part p0{};
        if (active) load_state(src, N, i, p0);
        bool need = true;
        if (a.prefilter) {
            // Pass 1 (no stores): particles that never reach the support and
            // carry no adjoint from later steps have a zero adjoint over the
            // whole segment; a wavefront of only such particles is skipped.
            const bool seeded = l.Z != 0.0 || l.R != 0.0 || l.Vz != 0.0 || l.Vr != 0.0 || l.Vt != 0.0 ||
                                l.Az != 0.0 || l.Ar != 0.0;
            bool touched = false;
            Part p = p0;
            for (int n = n0; n < nl; ++n) {
                push_core(PairField{a.Ezp, a.Erp, a.c[n]}, a.nz, a.nr, F, p);
                if (p.R <= 0.0) { p.R = fabs(p.R); p.Vr = fabs(p.Vr); p.Ar = fabs(p.Ar); }
                touched |= in_support(a, p);
            }
            need = __ballot(touched || seeded) != 0;
        }
        if (need) {
            // Pass 2: recompute with the tape of pre-push states.
            Part post = p0;
            for (int n = n0; n < nl; ++n) {
                tape_store(a.tape + (size_t)(n - n0) * kTape * N, N, i, post);
                push_core(PairField{a.Ezp, a.Erp, a.c[n]}, a.nz, a.nr, F, post);
                if (post.R <= 0.0) { post.R = fabs(post.R); post.Vr = fabs(post.Vr); post.Ar = fabs(post.Ar); }
            }
            // Reverse sweep, wave-uniform over the steps of the segment;
            // post is the state after step n (push and reflection).
            for (int n = n1 - 1; n >= n0; --n) {
                const bool live = n < nl;
                const double cn = a.c[n];
                double gc = 0.0;
                if (live) {
                    Part q;
                    tape_load(a.tape + (size_t)(n - n0) * kTape * N, N, i, q);
                    // loss of step n on the state after push and reflection
                    if (in_support(a, post)) {
                        double dz, dr;
                        const double m = mask_at(a, post.Z, post.R, dz, dr);
                        const double gl = a.g[n] * al * wq;
                        l.Z += gl * post.Vz * dz;
                        l.R += gl * post.Vz * dr;
                        l.Vz += gl * m;
                    }
                    if (post.R <= 0.0) {
                        // reflected (R' == 0 exactly): signs of the pushed Vr, Ar
                        Part q2 = q;
                        q2.Vz = 0.0; q2.Az = 0.0;
                        push_core(PairField{a.Ezp, a.Erp, cn}, a.nz, a.nr, F, q2);
                        if (q2.Vr < 0.0) l.Vr = -l.Vr;
                        if (q2.Ar < 0.0) l.Ar = -l.Ar;
                    }
                    double E2;
                    gc = push_adjoint(a, F, cn, q, l, E2);
                    // step back to the state before step n
                    const double az = post.Az;
                    post.Az = 2.0 * az - E2;
                    post.Vz = post.Vz - az - E2;
                    post.Z = q.Z; post.R = q.R; post.Vr = q.Vr; post.Vt = q.Vt; post.Ar = q.Ar;
                }
                if (__ballot(gc != 0.0)) {
                    const double w = wave_sum(gc);
                    if ((threadIdx.x & 63) == 0) lds_add_d(&sg[n - n0], w);
                }
            }
        }
        if (valid && al != 0.0 && j > 0) {
            gp(a.lam)[i] = l.Z; gp(a.lam + N)[i] = l.R; gp(a.lam + 2 * N)[i] = l.Vz;
            gp(a.lam + 3 * N)[i] = l.Vr; gp(a.lam + 4 * N)[i] = l.Vt; gp(a.lam + 5 * N)[i] = l.Az;
            gp(a.lam + 6 * N)[i] = l.Ar;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n1 - n0; t += kWBlock)
        if (sg[t] != 0.0) global_add(&a.acc[n0 + t], sg[t]);
}

__global__ void gather_species_kernel(double *ck0, double *wq, const DevState *stp, int s, long long n,
                                      long long off, long long N)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *const *F = stp->f[s];
    for (int f = 0; f < kNS; ++f) ck0[(size_t)f * N + off + i] = F[f][i];
    wq[off + i] = F[7][i];
}

}  // namespace

struct ipm_window {
    WinArgs a{};
    double *ck = nullptr, *wq = nullptr, *tape = nullptr, *lam = nullptr;
    Pair2 *ezp = nullptr, *erp = nullptr;
    double *mu = nullptr, *tbuf = nullptr;   // tbuf: c, acc, g (3 * T)
    int *death = nullptr;
    bool recorded = false;
    bool ck_valid = false;
    std::vector<double> c_ck;                // coefficients the checkpoints were made with
    hipEvent_t ev[4] = {};
    float fwd_ms = 0.f, bwd_ms = 0.f;
};

int window_release(ipm_ctx *c)
{
    if (!c || !c->win) return IPM_OK;
    ipm_window *w = c->win;
    void *ptrs[] = {w->ck, w->wq, w->tape, w->lam, w->ezp, w->erp, w->mu, w->tbuf, w->death};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : w->ev)
        if (e) (void)hipEventDestroy(e);
    delete w;
    c->win = nullptr;
    return IPM_OK;
}

static int window_grid(ipm_ctx *c, long long N)
{
    const long long need = (N + kWBlock - 1) / kWBlock;
    return (int)std::max<long long>(1, std::min<long long>(need, (long long)c->num_cu * 8));
}

static int window_run_forward(ipm_ctx *c, const double *coef, int record)
{
    ipm_window *w = c->win;
    WinArgs &a = w->a;
    HIPCHK(hipMemcpyAsync(w->tbuf, coef, sizeof(double) * a.T, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(w->tbuf + a.T, 0, sizeof(double) * a.T, c->stream));
    a.c = w->tbuf;
    a.acc = w->tbuf + a.T;
    const int grid = window_grid(c, a.N);
    HIPCHK(hipEventRecord(w->ev[0], c->stream));
    for (int j = 0; j < a.nseg; ++j)
        hipLaunchKernelGGL(win_forward_kernel, dim3(grid), dim3(kWBlock), 0, c->stream, a, j, record);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(w->ev[1], c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventElapsedTime(&w->fwd_ms, w->ev[0], w->ev[1]));
    if (record) w->recorded = true;
    w->c_ck.assign(coef, coef + a.T);
    w->ck_valid = true;
    return IPM_OK;
}

extern "C" {

int ipm_window_create(ipm_ctx *c, int nsteps, int interval, const double *Ez0, const double *Er0,
                      const double *Ez1, const double *Er1, const double *mu, const double *alpha)
{
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    if (nsteps < 1 || interval < 1 || interval > kKMax)
        return fail(IPM_EINVAL, "window: need nsteps >= 1 and 1 <= interval <= %d (got %d, %d)", kKMax, nsteps,
                    interval);
    if (!Ez0 || !Er0 || !Ez1 || !Er1 || !mu || !alpha) return fail(IPM_EINVAL, "window: null array");
    window_release(c);
    DevCounters ctr;
    rc = read_counters(c, &ctr);
    if (rc) return rc;
    ipm_window *w = new ipm_window();
    c->win = w;
    WinArgs &a = w->a;
    a.nz = c->nz; a.nr = c->nr; a.ns = c->ns; a.T = nsteps; a.K = interval;
    a.nseg = (nsteps + interval - 1) / interval;
    long long N = 0;
    for (int s = 0; s < c->ns; ++s) {
        a.off[s] = N;
        N += (long long)ctr.n[s];
        a.efactor[s] = c->h.efactor[s];
        a.alpha[s] = alpha[s];
    }
    for (int s = c->ns; s <= IPM_NS_MAX; ++s) a.off[s] = N;
    a.N = N;
    const size_t NN = (size_t)std::max<long long>(N, 1);
    const int nz = c->nz, nr = c->nr;
    HIPCHK(hipMalloc(&w->ck, sizeof(double) * kNS * NN * (a.nseg + 1)));
    HIPCHK(hipMalloc(&w->wq, sizeof(double) * NN));
    HIPCHK(hipMalloc(&w->death, sizeof(int) * NN));
    HIPCHK(hipMalloc(&w->tape, sizeof(double) * kTape * NN * interval));
    HIPCHK(hipMalloc(&w->lam, sizeof(double) * kNS * NN));
    HIPCHK(hipMalloc(&w->ezp, sizeof(Pair2) * (size_t)(nz + 1) * nr));
    HIPCHK(hipMalloc(&w->erp, sizeof(Pair2) * (size_t)nz * (nr + 1)));
    HIPCHK(hipMalloc(&w->mu, sizeof(double) * (size_t)nz * nr));
    HIPCHK(hipMalloc(&w->tbuf, sizeof(double) * 3 * nsteps));
    for (hipEvent_t &e : w->ev) HIPCHK(hipEventCreate(&e));
    // interleaved (E0, E1) pairs
    {
        std::vector<Pair2> pz((size_t)(nz + 1) * nr), pr((size_t)nz * (nr + 1));
        for (size_t k = 0; k < pz.size(); ++k) pz[k] = Pair2{Ez0[k], Ez1[k]};
        for (size_t k = 0; k < pr.size(); ++k) pr[k] = Pair2{Er0[k], Er1[k]};
        HIPCHK(hipMemcpy(w->ezp, pz.data(), sizeof(Pair2) * pz.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(w->erp, pr.data(), sizeof(Pair2) * pr.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(w->mu, mu, sizeof(double) * (size_t)nz * nr, hipMemcpyHostToDevice));
    // support box of mu (1-based node indices imin..imax, jmin..jmax): the
    // gather at (Z, R) reads nodes ceil(Z), ceil(Z)+1 (after the guards), so
    // it meets the box only if imin - 1 <= ceil(Z) <= imax.
    int imin = nz + 1, imax = 0, jmin = nr + 1, jmax = 0;
    for (int jr = 1; jr <= nr; ++jr)
        for (int iz = 1; iz <= nz; ++iz)
            if (mu[(size_t)(iz - 1) + (size_t)(jr - 1) * nz] != 0.0) {
                imin = std::min(imin, iz); imax = std::max(imax, iz);
                jmin = std::min(jmin, jr); jmax = std::max(jmax, jr);
            }
    if (imax == 0) {
        a.zlo = a.zhi = a.rlo = a.rhi = 0.0;     // empty support: nothing passes
    } else {
        // guards clamp ceil(Z) into [1, nz-1]; widen the box where they apply
        a.zlo = (imin <= 2) ? -INFINITY : (double)(imin - 2);
        a.zhi = (imax >= nz - 1) ? INFINITY : (double)imax;
        a.rlo = (jmin <= 2) ? -INFINITY : (double)(jmin - 2);
        a.rhi = (jmax >= nr - 1) ? INFINITY : (double)jmax;
    }
    // The skip test costs one extra recompute; it pays when the support is a
    // small part of the grid (a flux through a few planes), not otherwise.
    {
        const double fz = (std::min(a.zhi, (double)(nz - 1)) - std::max(a.zlo, 0.0)) / (double)(nz - 1);
        const double fr = (std::min(a.rhi, (double)(nr - 1)) - std::max(a.rlo, 0.0)) / (double)(nr - 1);
        a.prefilter = (imax > 0 && fz * fr < 0.5) ? 1 : 0;
    }
    a.Ezp = w->ezp; a.Erp = w->erp; a.mu = w->mu;
    a.cell_material = c->cell_material; a.mat_type = c->mat_type; a.ndiel = c->h.ndiel;
    a.ck = w->ck; a.wq = w->wq; a.death = w->death; a.tape = w->tape; a.lam = w->lam;
    a.c = w->tbuf; a.acc = w->tbuf + nsteps; a.g = w->tbuf + 2 * nsteps;
    for (int s = 0; s < c->ns; ++s) {
        const long long n = (long long)ctr.n[s];
        if (n == 0) continue;
        hipLaunchKernelGGL(gather_species_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           w->ck, w->wq, c->d_state, s, n, a.off[s], N);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(w->death, 0xff, sizeof(int) * NN, c->stream));   // -1 until recorded
    HIPCHK(hipStreamSynchronize(c->stream));
    return IPM_OK;
}

int ipm_window_forward(ipm_ctx *c, const double *coef, int record, double *loss)
{
    if (!c || !c->win) return fail(IPM_EINVAL, "window: no window (call ipm_window_create)");
    if (!coef) return fail(IPM_EINVAL, "window: null coefficients");
    ipm_window *w = c->win;
    if (!record && !w->recorded)
        return fail(IPM_EINVAL, "window: replay needs a recording forward first (record = 1)");
    int rc = window_run_forward(c, coef, record ? 1 : 0);
    if (rc) return rc;
    if (loss) HIPCHK(hipMemcpy(loss, w->a.acc, sizeof(double) * w->a.T, hipMemcpyDeviceToHost));
    return IPM_OK;
}

int ipm_window_backward(ipm_ctx *c, const double *coef, const double *gloss, double *dcoef)
{
    if (!c || !c->win) return fail(IPM_EINVAL, "window: no window (call ipm_window_create)");
    if (!coef || !gloss || !dcoef) return fail(IPM_EINVAL, "window: null array");
    ipm_window *w = c->win;
    WinArgs &a = w->a;
    if (!w->recorded) return fail(IPM_EINVAL, "window: the adjoint needs a recording forward first");
    // checkpoints made with other coefficients are rebuilt by a replay
    if (!w->ck_valid || memcmp(w->c_ck.data(), coef, sizeof(double) * a.T) != 0) {
        int rc = window_run_forward(c, coef, 0);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(w->tbuf, coef, sizeof(double) * a.T, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(w->tbuf + 2 * a.T, gloss, sizeof(double) * a.T, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(w->tbuf + a.T, 0, sizeof(double) * a.T, c->stream));
    const int grid = window_grid(c, a.N);
    HIPCHK(hipEventRecord(w->ev[2], c->stream));
    for (int j = a.nseg - 1; j >= 0; --j)
        hipLaunchKernelGGL(win_adjoint_kernel, dim3(grid), dim3(kWBlock), 0, c->stream, a, j);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(w->ev[3], c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventElapsedTime(&w->bwd_ms, w->ev[2], w->ev[3]));
    HIPCHK(hipMemcpy(dcoef, a.acc, sizeof(double) * a.T, hipMemcpyDeviceToHost));
    return IPM_OK;
}

int ipm_window_info(ipm_ctx *c, int64_t *nparticles, int *nsegments, float *forward_ms, float *adjoint_ms)
{
    if (!c || !c->win) return fail(IPM_EINVAL, "window: no window (call ipm_window_create)");
    if (nparticles) *nparticles = c->win->a.N;
    if (nsegments) *nsegments = c->win->a.nseg;
    if (forward_ms) *forward_ms = c->win->fwd_ms;
    if (adjoint_ms) *adjoint_ms = c->win->bwd_ms;
    return IPM_OK;
}

int ipm_window_download(ipm_ctx *c, int checkpoint, double *state, int *absorbed_step)
{
    if (!c || !c->win) return fail(IPM_EINVAL, "window: no window (call ipm_window_create)");
    ipm_window *w = c->win;
    if (checkpoint < 0 || checkpoint > w->a.nseg) return fail(IPM_EINVAL, "window: checkpoint %d out of range", checkpoint);
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t N = (size_t)w->a.N;
    if (state && N)
        HIPCHK(hipMemcpy(state, w->ck + (size_t)checkpoint * kNS * N, sizeof(double) * kNS * N,
                         hipMemcpyDeviceToHost));
    if (absorbed_step && N) HIPCHK(hipMemcpy(absorbed_step, w->death, sizeof(int) * N, hipMemcpyDeviceToHost));
    return IPM_OK;
}

int ipm_window_destroy(ipm_ctx *c)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    if (c->stream) HIPCHK(hipStreamSynchronize(c->stream));
    return window_release(c);
}

}  // extern "C"
