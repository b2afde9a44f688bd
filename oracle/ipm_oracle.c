/*
 * CPU oracle for the iPM2D particle hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Statement-by-statement C restatement of the reference routines cited at
 * each function.  Compiled with -ffp-contract=off so that every product and
 * sum is rounded exactly as the Fortran source orders it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this code,
 * always as the checker; the product never links or calls it.
 *
 * Pinning (what checks this restatement, beyond retyping the reference):
 *   ora_philox4x32_10     Random123 known-answer vectors
 *   ora_push              closed forms (E=0 straight Cartesian line, uniform
 *                         Ez quadratic z, linear fields), invariants (R*Vt,
 *                         Vr^2+Vt^2), axis branch, random-field gathers
 *                         bounded by the bracketing faces' values
 *   ora_absorb            rule table on hand-built particles, exact
 *   ora_deposit_*         charge conservation, on-node known answers,
 *                         uniform density vs analytic node volumes
 *   ora_node_volumes      closed-form cylindrical shell volumes
 *   ora_solver_coe/_source  uniform medium = the cylindrical five-point
 *                         Laplacian; a quadratic potential's stencil gives
 *                         (2a + 4b)(1 + chi) eps; flux symmetry A(i+1,j) =
 *                         C(i,j), (j-1) E(i,j) = j D(i,j+1) for any medium;
 *                         hand-worked axis and aliased-cell values; Chi = 0
 *                         form = form without Chi; source closed form
 *   ora_set_e             linear potentials give the uniform field exactly
 *                         on every face including the ghost-fed ones;
 *                         numpy restatement agrees bit for bit
 *   ora_moments           charge conservation, macro-particle count, uniform
 *                         velocity and speed known answers, implicit current
 *                         equal to the explicit one in a zero field
 *   ora_mcc               selection ratio 1-exp(-nu dt), channel fractions
 *                         n sigma v, Vahedi/isotropic cos-chi CDFs,
 *                         Maxwellian mean energy, energy bookkeeping of
 *                         excitation/ionization, momentum of elastic events;
 *                         through ora_collide_given: hand-worked 111/112/113
 *                         velocities (with the Vy <- Vz term), charge
 *                         exchange = target velocity, CM speed conserved by
 *                         101/11 with isotropic/Vahedi mean deflection,
 *                         the ionization secondary's direction and inverted
 *                         CreateEnergy draw, dissociation product species;
 *                         Langevin fraction 1/beta_max^2 and the
 *                         sub-Threshold(2) shortcut; gases in loop order
 *   ora_probability_hard_sphere  sigma = 2 pi ((r_g+r_s)/2)^2, v(E) shape,
 *                         the 0.5/1.0 split (equal 101/103 event counts)
 *   ora_sigma_* / ora_probability_*  closed forms on synthetic tables,
 *                         including sigma linear in the log-energy index
 *                         (the dble(i) weight) and the Emin coordinate of
 *                         raw energies below the grid
 *   ora_recombine         hand-computed single-cell budgets (constant and
 *                         temperature forms), the visiting order's two
 *                         outcomes at their closed-form frequencies, cells
 *                         independent of their interleaving, passes of
 *                         later reactions skipping earlier deletions;
 *                         ora_recomb_density: closed-form densities and
 *                         mean energies
 *   ora_split_merge       hand-computed splits (copy counts, remainders,
 *                         total weight), merge keep fraction = WQ/ref within
 *                         binomial bounds, expected weight conserved, the
 *                         copies' Vt differing from the parent's by the XOR
 *                         tag only
 * Parity unpinned: the reference's own random stream (Fortran
 * random_number, unseeded) -- by construction no check can reproduce it.
 */
#include "ipm_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

/* R/src/base/Constants.F90:3-15 */
#define C_PI 3.141592653589793238
#define C_KB 1.3807e-23
#define C_EXP 2.71828
#define C_E 1.6022e-19
#define C_EPS0 8.8542e-12
#define C_JTOEV 1.6022e-19
#define C_MINREAL 1.0e-15

/* ---------------------------------------------------------------- RNG */
void ora_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Draw `draw` of the stream whose 96-bit counter prefix is (c0, c1, c2):
 * block {c0, c1, c2, draw/2}, key {seed_lo + step*0x9E3779B9,
 * seed_hi ^ tag*0x85EBCA6B}; the top 52 bits of words (1,0) for an even
 * draw, of words (3,2) for an odd one -> (k + 1/2) 2^-52 in (0,1). */
double ora_uniform(uint64_t seed, uint32_t step, uint32_t tag, uint32_t c0, uint32_t c1, uint32_t c2,
                   uint32_t draw)
{
    uint32_t ctr[4] = {c0, c1, c2, draw >> 1};
    uint32_t key[2] = {(uint32_t)seed + step * 0x9E3779B9u, (uint32_t)(seed >> 32) ^ (tag * 0x85EBCA6Bu)};
    uint32_t o[4];
    ora_philox4x32_10(ctr, key, o);
    uint64_t x = (draw & 1u) ? ((((uint64_t)o[3]) << 32) | o[2]) >> 12    /* 52 bits */
                             : ((((uint64_t)o[1]) << 32) | o[0]) >> 12;
    return ((double)x + 0.5) * 0x1p-52;                    /* in (0,1) */
}

static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static uint64_t dbits(double d)
{
    uint64_t u;
    memcpy(&u, &d, sizeof u);
    return u;
}

/* 96-bit stream prefix folded from a particle's phase-space state: the
 * stream follows the particle, not its storage slot. */
void ora_state_key(double Z, double R, double Vz, double Vr, double Vt, uint32_t out[3])
{
    uint64_t z = dbits(Z), r = dbits(R), vz = dbits(Vz), vr = dbits(Vr), vt = dbits(Vt);
    out[0] = (uint32_t)z ^ (uint32_t)(vr >> 32) ^ rotl32((uint32_t)(vt >> 32), 13);
    out[1] = (uint32_t)r ^ (uint32_t)(vz >> 32) ^ rotl32((uint32_t)vt, 7);
    out[2] = (uint32_t)vz ^ rotl32((uint32_t)vr, 11) ^ rotl32((uint32_t)(z >> 32), 3) ^ rotl32((uint32_t)(r >> 32), 19);
}

static double flip_low_bit(double d)
{
    uint64_t u = dbits(d) ^ 1u;
    double r;
    memcpy(&r, &u, sizeof r);
    return r;
}

typedef struct { uint64_t seed; uint32_t step, tag, c0, c1, c2, k; } rng_t;
static double rnext(rng_t *r) { return ora_uniform(r->seed, r->step, r->tag, r->c0, r->c1, r->c2, r->k++); }

/* ---------------------------------------------------------- rezoning */
#define ORA_MERGE_TAG 0x4D520000u

/* CalRW, R/src/pic/particle/ParticleBundle.f90:393-402 */
static double cal_rw(double weight, double alpha, double grid_r, double pos_r)
{
    return weight * ((1.0 - alpha) * pos_r / grid_r + alpha);
}

/* split_particle_weight (R/src/base/Rezoning.F90:73-92) then
 * merge_particle_random (Rezoning.F90:53-70) on one bundle of n particles in
 * arrays of capacity cap.  The uniform of the merge is drawn from the
 * particle's state-keyed stream, tag ORA_MERGE_TAG | s, draw 0, in place of
 * Fortran's random_number.  Copy j of a split carries the parent's Vt bit
 * pattern XOR j.  A particle whose reference weight is not positive is
 * skipped (the reference divides by it).  Returns -1 when the copies would
 * exceed cap (nothing is changed then), 0 otherwise.  stats: copies
 * appended, merge candidates kept, merge candidates deleted. */
int ora_split_merge(int64_t *n, double *P[8], int64_t cap, int mode, double weight, double alpha,
                    double grid_r, double merge_limit, uint64_t seed, uint32_t step, int s, int64_t stats[3])
{
    double *Z = P[0], *R = P[1], *Vt = P[4], *WQ = P[7];
    stats[0] = stats[1] = stats[2] = 0;
    if (mode & 1) {
        int64_t need = 0;
        for (int64_t i = 0; i < *n; ++i) {
            double ref = cal_rw(weight, alpha, grid_r, R[i]);
            if (ref > 0.0 && WQ[i] / ref >= 2.0) {
                double ratio = WQ[i] / ref;
                if (!(ratio - 1.0 <= (double)(cap - *n))) return -1;
                need += (int64_t)ratio - 1;
            }
        }
        if (need > cap - *n) return -1;
        const int64_t npar = *n;
        for (int64_t i = 0; i < npar; ++i) {
            double ref = cal_rw(weight, alpha, grid_r, R[i]);
            if (!(ref > 0.0)) continue;
            int64_t ncr = (int64_t)(WQ[i] / ref);
            if (ncr > 1) {
                for (int64_t j = 1; j <= ncr - 1; ++j) {
                    WQ[i] = WQ[i] - ref;
                    int64_t q = *n;
                    for (int f = 0; f < 8; ++f) P[f][q] = P[f][i];
                    uint64_t b;
                    memcpy(&b, &Vt[i], 8);
                    b ^= (uint64_t)j;
                    memcpy(&Vt[q], &b, 8);
                    WQ[q] = ref;
                    *n += 1;
                    stats[0] += 1;
                }
            }
        }
    }
    if (mode & 2) {
        for (int64_t i = *n - 1; i >= 0; --i) {
            double ref = cal_rw(weight, alpha, grid_r, R[i]);
            if (!(ref > 0.0)) continue;
            if (WQ[i] < merge_limit * ref) {
                uint32_t k[3];
                ora_state_key(Z[i], R[i], P[2][i], P[3][i], Vt[i], k);
                double u = ora_uniform(seed, step, ORA_MERGE_TAG | (uint32_t)s, k[0], k[1], k[2], 0);
                if (u < WQ[i] / ref) {
                    WQ[i] = ref;
                    stats[1] += 1;
                } else {
                    for (int f = 0; f < 8; ++f) P[f][i] = P[f][*n - 1];
                    *n -= 1;
                    stats[2] += 1;
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------ recombination */
#define ORA_RECOMB_TAG 0x52430000u

static int64_t recomb_cell(int nz, int nr, double z, double r)
{
    int iz = (int)ceil(z), ir = (int)ceil(r);
    iz = iz < 1 ? 1 : (iz > nz - 1 ? nz - 1 : iz);
    ir = ir < 1 ? 1 : (ir > nr - 1 ? nr - 1 : ir);
    return (int64_t)(iz - 1) + (int64_t)(ir - 1) * (nz - 1);
}

/* Volume of radial cell ir, initRecombination2D (Recombination.F90:133-137). */
static double recomb_volume(double dx, int ir)
{
    return C_PI * dx * dx * dx * (double)(ir * ir - (ir - 1) * (ir - 1));
}

/* Cell number density and, when want_t, inverse mean energy (eV) of one
 * bundle: executingRecombination2D, Recombination.F90:175-206.  dens and
 * inv_t hold (nz-1)*(nr-1) cells. */
void ora_recomb_density(int nz, int nr, double dx, double dt, double mass, int want_t, int64_t n,
                        double *const P[8], double *dens, double *inv_t)
{
    const int64_t ncell = (int64_t)(nz - 1) * (nr - 1);
    const double vf = dx / dt;
    for (int64_t c = 0; c < ncell; ++c) {
        dens[c] = 0.0;
        if (want_t) inv_t[c] = 0.0;
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t c = recomb_cell(nz, nr, P[0][i], P[1][i]);
        dens[c] = dens[c] + P[7][i];
        if (want_t) {
            double energy = P[2][i] * P[2][i] + P[3][i] * P[3][i] + P[4][i] * P[4][i];
            inv_t[c] = inv_t[c] + energy * P[7][i];
        }
    }
    for (int64_t c = 0; c < ncell; ++c) {
        int ir = (int)(c / (nz - 1)) + 1;
        double inv_vol = 1.0 / recomb_volume(dx, ir);
        dens[c] = dens[c] * inv_vol;
        if (want_t) {
            inv_t[c] = inv_t[c] * inv_vol * 0.5 * mass * vf * vf;
            if (dens[c] > 0.0 && inv_t[c] > 0.0)
                inv_t[c] = 1.0 / (inv_t[c] / C_JTOEV / dens[c]);
            else
                inv_t[c] = 0.0;
        }
    }
}

/* deletingParticles, Recombination.F90:246-273, visiting i = n-1 .. 0 and
 * skipping particles marked by earlier passes of the same call; deletion is
 * a mark (Z = -1e300). */
static int64_t recomb_delete(int nz, int nr, int64_t n, double *const P[8], double *recb, uint64_t seed,
                             uint32_t step, uint32_t tag)
{
    int64_t cnt = 0;
    for (int64_t i = n - 1; i >= 0; --i) {
        if (P[0][i] < -1.0e299) continue;
        int64_t c = recomb_cell(nz, nr, P[0][i], P[1][i]);
        if (recb[c] > P[7][i]) {
            recb[c] = recb[c] - P[7][i];
            P[0][i] = -1.0e300;
            ++cnt;
        } else if (recb[c] > 0.0) {
            uint32_t k[3];
            ora_state_key(P[0][i], P[1][i], P[2][i], P[3][i], P[4][i], k);
            double u = ora_uniform(seed, step, tag, k[0], k[1], k[2], 0);
            if (u < recb[c] / P[7][i]) {
                recb[c] = 0.0;
                P[0][i] = -1.0e300;
                ++cnt;
            }
        }
    }
    return cnt;
}

/* executingRecombination2D, Recombination.F90:164-243, on ns bundles of
 * n[s] particles (P holds 8 arrays per species, species-major).  reactions
 * is recb_reaction(nreaction, 3) in Fortran order with 0-based species;
 * type 0 uses a constant rate, type 1 the temperature form.  The draw of
 * reaction i's pass over species x (role 0 = first, 1 = second species) is
 * the particle's state-keyed uniform under tag ORA_RECOMB_TAG | i<<4 |
 * role<<3 | x.  Survivors are compacted, order kept.  Returns -1 on a bad
 * reaction. */
int ora_recombine(int nz, int nr, double dx, double dt, int ns, int64_t *n, double *const *P, const double *mass,
                  int nreaction, const int *reactions, const double *rate, double time_step, uint64_t seed,
                  uint32_t step, int64_t *deleted)
{
    const int64_t ncell = (int64_t)(nz - 1) * (nr - 1);
    int involved[16] = {0}, want_t[16] = {0};
    if (ns > 16) return -1;
    for (int i = 0; i < nreaction; ++i) {
        int a = reactions[i], b = reactions[nreaction + i], t = reactions[2 * nreaction + i];
        if (a < 0 || a >= ns || b < 0 || b >= ns || (t != 0 && t != 1)) return -1;
        involved[a] = involved[b] = 1;
        if (t == 1) want_t[a] = want_t[b] = 1;
    }
    double *dens = (double *)calloc((size_t)(ns * ncell), sizeof(double));
    double *inv_t = (double *)calloc((size_t)(ns * ncell), sizeof(double));
    double *recb_neg = (double *)malloc(sizeof(double) * (size_t)ncell);
    double *recb_pos = (double *)malloc(sizeof(double) * (size_t)ncell);
    for (int s = 0; s < ns; ++s) {
        deleted[s] = 0;
        if (involved[s])
            ora_recomb_density(nz, nr, dx, dt, mass[s], want_t[s], n[s], P + 8 * s, dens + s * ncell,
                               inv_t + s * ncell);
    }
    for (int i = 0; i < nreaction; ++i) {
        int a = reactions[i], b = reactions[nreaction + i], t = reactions[2 * nreaction + i];
        const double *da = dens + a * ncell, *db = dens + b * ncell, *ta = inv_t + a * ncell, *tb = inv_t + b * ncell;
        for (int64_t c = 0; c < ncell; ++c) {
            double vol = recomb_volume(dx, (int)(c / (nz - 1)) + 1);
            if (t == 0)
                recb_neg[c] = da[c] * db[c] * rate[i] * time_step * vol;
            else
                recb_neg[c] = da[c] * db[c] * time_step * vol * rate[i] * sqrt(ta[c]) * tb[c];
            recb_pos[c] = recb_neg[c];
        }
        deleted[a] += recomb_delete(nz, nr, n[a], P + 8 * a, recb_neg, seed, step,
                                    ORA_RECOMB_TAG | ((uint32_t)i << 4) | (uint32_t)a);
        deleted[b] += recomb_delete(nz, nr, n[b], P + 8 * b, recb_pos, seed, step,
                                    ORA_RECOMB_TAG | ((uint32_t)i << 4) | (1u << 3) | (uint32_t)b);
    }
    for (int s = 0; s < ns; ++s) {
        if (!involved[s]) continue;
        double *const *Q = P + 8 * s;
        int64_t j = 0;
        for (int64_t i = 0; i < n[s]; ++i) {
            if (Q[0][i] < -1.0e299) continue;
            for (int f = 0; f < 8; ++f) Q[f][j] = Q[f][i];
            ++j;
        }
        n[s] = j;
    }
    free(dens);
    free(inv_t);
    free(recb_neg);
    free(recb_pos);
    return 0;
}

/* ------------------------------------------------------- node volumes */
void ora_node_volumes(int nz, int nr, double dx, const double *factor, double *vol, double *inv_vol)
{
    for (int j = 1; j <= nr; ++j)
        for (int i = 1; i <= nz; ++i) {
            size_t id = (size_t)(i - 1) + (size_t)(j - 1) * nz;
            double f = factor[id], v;
            if (j == 1)
                v = f * 1.0 / 3.0 * C_PI * dx * dx * dx;
            else if (j == nr)
                v = f * 1.0 / 3.0 * ((double)(j - 2) + 2.0 * (double)(j - 1)) * C_PI * dx * dx * dx;
            else
                v = f * 1.0 / 3.0 * ((double)j * ((double)(j - 1) + (double)j) -
                                     (double)(j - 2) * ((double)(j - 2) + (double)(j - 1))) * C_PI * dx * dx * dx;
            vol[id] = v;
            inv_vol[id] = (f > 1.0e-6) ? 1.0 / v : 0.0;
        }
}

/* --------------------------------------------------------------- push */
void ora_push(int nz, int nr, double dx, double dt, double charge, double mass,
              const double *Ez, const double *Er, int64_t n,
              double *Z, double *R, double *Vz, double *Vr, double *Vt, double *Az, double *Ar)
{
    (void)nr;
    const double EFactor = charge / mass * dt / (dx / dt) * 0.5;
    const size_t ldz = (size_t)nz + 1;   /* Ez(0:nz, 1:nr) */
    const size_t ldr = (size_t)nz;       /* Er(1:nz, 0:nr) */
#define EZ(i, j) Ez[(size_t)(i) + ((size_t)(j) - 1) * ldz]
#define ER(i, j) Er[((size_t)(i) - 1) + (size_t)(j) * ldr]
    for (int64_t p = 0; p < n; ++p) {
        double z = Z[p], r = R[p];
        int N1 = (int)ceil(z);
        double Z1 = (double)N1 - z, Z2 = 1.0 - Z1;
        int N1H = (int)(z + 0.5);
        double Z1H = (double)N1H + 0.5 - z, Z2H = 1.0 - Z1H;
        int N2 = (int)ceil(r);
        double R1 = (double)N2 - r, R2 = 1.0 - R1;
        int N2H = (int)(r + 0.5);
        double R1H = (double)N2H + 0.5 - r, R2H = 1.0 - R1H;
        /* R == 0 exactly gives N2 == 0 (a row outside Ez); its weight R1 is 0 */
        int N2e = N2 < 1 ? 1 : N2;

        double E1 = ER(N1, N2H) * Z1 * R1H + ER(N1 + 1, N2H) * Z2 * R1H +
                    ER(N1, N2H + 1) * Z1 * R2H + ER(N1 + 1, N2H + 1) * Z2 * R2H;
        double E2 = EZ(N1H, N2e) * Z1H * R1 + EZ(N1H + 1, N2e) * Z2H * R1 +
                    EZ(N1H, N2 + 1) * Z1H * R2 + EZ(N1H + 1, N2 + 1) * Z2H * R2;
        E1 = E1 * EFactor;
        E2 = E2 * EFactor;

        double vz = Vz[p] + E2;
        z = z + E2;
        double az = (Az[p] + E2) * 0.5;
        double vr = Vr[p] + E1;
        r = r + E1;
        double ar = (Ar[p] + E1) * 0.5;

        vz = vz + az;
        z = z + vz;

        double P_R = r;
        double Vx = vr + ar;
        double Vy = Vt[p];
        double X = P_R + Vx;
        double Y = Vy;
        P_R = sqrt(X * X + Y * Y);
        r = P_R;
        double vt;
        if (r < C_MINREAL) {
            vr = sqrt(Vx * Vx + Vy * Vy);
            vt = C_MINREAL;
            ar = fabs(ar);
        } else {
            double SinA = Y / P_R;
            double CosA = X / P_R;
            vr = Vx * CosA + Vy * SinA;
            vt = -Vx * SinA + Vy * CosA;
            ar = ar * CosA;
        }
        Z[p] = z; R[p] = r; Vz[p] = vz; Vr[p] = vr; Vt[p] = vt; Az[p] = az; Ar[p] = ar;
    }
#undef EZ
#undef ER
}

/* --------------------------------------------------------- absorption */
#define MAT_METAL 1
#define MAT_DIEL 2
#define EDGE_V2D 1
#define EDGE_D2V 3

static void delone(double *P[8], int64_t *n, int64_t i)
{
    for (int f = 0; f < 8; ++f) P[f][i] = P[f][*n - 1];
    *n -= 1;
}

void ora_absorb(const ora_geom *g, double charge, int64_t *n, double *P[8],
                double *metal_charge, double *sigma)
{
    const int nz = g->nz, nr = g->nr;
#define CM(i, j) g->cell_material[((size_t)(i) - 1) + ((size_t)(j) - 1) * (nz - 1)]
#define HE(i, j) g->hedge_type[((size_t)(i) - 1) + ((size_t)(j) - 1) * nz]
#define VE(i, j) g->vedge_type[((size_t)(i) - 1) + ((size_t)(j) - 1) * (nz - 1)]
#define SG(i, j) sigma[((size_t)(i) - 1) + ((size_t)(j) - 1) * nz]
    for (int64_t i = *n - 1; i >= 0; --i) {
        double *Z = P[0], *R = P[1], *Vr = P[3], *Ar = P[6], *WQ = P[7];
        if (Z[i] <= 0.0 || Z[i] >= (double)(nz - 1) || R[i] >= (double)(nr - 1)) {
            delone(P, n, i);
        } else if (R[i] <= 0.0) {
            R[i] = fabs(R[i]);
            Vr[i] = fabs(Vr[i]);
            Ar[i] = fabs(Ar[i]);
        } else {
            int NzL = (int)ceil(Z[i]);
            int NrL = (int)ceil(R[i]);
            int m = CM(NzL, NrL);
            int t = g->material_type[m + g->ndiel];
            if (t == MAT_METAL) {
                metal_charge[m - 1] = metal_charge[m - 1] + WQ[i] * charge;
                delone(P, n, i);
            } else if (t == MAT_DIEL) {
                int h0 = HE(NzL, NrL), h1 = HE(NzL + 1, NrL);
                int v0 = VE(NzL, NrL), v1 = VE(NzL, NrL + 1);
                if (h0 == EDGE_V2D || h0 == EDGE_D2V) {
                    double RL = (double)NrL - R[i], RU = 1.0 - RL;
                    SG(NzL, NrL) = SG(NzL, NrL) + RL * WQ[i] * charge;
                    SG(NzL, NrL + 1) = SG(NzL, NrL + 1) + RU * WQ[i] * charge;
                } else if (h1 == EDGE_V2D || h1 == EDGE_D2V) {
                    double RL = (double)NrL - R[i], RU = 1.0 - RL;
                    SG(NzL + 1, NrL) = SG(NzL + 1, NrL) + RL * WQ[i] * charge;
                    SG(NzL + 1, NrL + 1) = SG(NzL + 1, NrL + 1) + RU * WQ[i] * charge;
                } else if (v0 == EDGE_V2D || v0 == EDGE_D2V) {
                    double ZL = (double)NzL - Z[i], ZU = 1.0 - ZL;
                    SG(NzL, NrL) = SG(NzL, NrL) + ZL * WQ[i] * charge;
                    SG(NzL + 1, NrL) = SG(NzL + 1, NrL) + ZU * WQ[i] * charge;
                } else if (v1 == EDGE_V2D || v1 == EDGE_D2V) {
                    double ZL = (double)NzL - Z[i], ZU = 1.0 - ZL;
                    SG(NzL, NrL + 1) = SG(NzL, NrL + 1) + ZL * WQ[i] * charge;
                    SG(NzL + 1, NrL + 1) = SG(NzL + 1, NrL + 1) + ZU * WQ[i] * charge;
                }
                delone(P, n, i);
            }
        }
    }
#undef CM
#undef HE
#undef VE
#undef SG
}

/* ------------------------------------------------------------ deposit */
void ora_deposit_raw(int nz, int nr, int64_t n, const double *Z, const double *R,
                     const double *WQ, double *rho_raw)
{
    (void)nr;
#define RH(i, j) rho_raw[((size_t)(i) - 1) + ((size_t)(j) - 1) * nz]
    for (int64_t p = 0; p < n; ++p) {
        int NzL = (int)ceil(Z[p]);
        int NzU = NzL + 1;
        double ZL = (double)NzL - Z[p];
        double ZU = 1.0 - ZL;
        int NrL = (int)ceil(R[p]);
        int NrU = NrL + 1;
        double RL = (double)NrL - R[p];
        double RU = 1.0 - RL;
        /* R == 0 exactly: row NrL = 0 lies outside Rho and its weight RL is 0 */
        if (NrL >= 1) RH(NzL, NrL) = RH(NzL, NrL) + ZL * RL * WQ[p];
        RH(NzL, NrU) = RH(NzL, NrU) + ZL * RU * WQ[p];
        if (NrL >= 1) RH(NzU, NrL) = RH(NzU, NrL) + ZU * RL * WQ[p];
        RH(NzU, NrU) = RH(NzU, NrU) + ZU * RU * WQ[p];
    }
#undef RH
}

void ora_deposit_finish(int nz, int nr, double dt, double charge, double mass,
                        const double *rho_raw, const double *inv_vol, double *rho, double *chi)
{
    const double ChiFactor = 0.5 * charge / mass * dt * dt / C_EPS0;
    size_t nn = (size_t)nz * nr;
    for (size_t k = 0; k < nn; ++k) {
        double v = rho_raw[k] * inv_vol[k];
        v = v * charge;
        rho[k] = v;
        chi[k] = v * ChiFactor;
    }
}

/* SetE: FT%SetE(FG), R/src/pic/field/FieldSource.f90:95-120, after the
 * potential's ghost layers are filled by linear extrapolation, the
 * reference's default for every side of a one-rank domain
 * (R/src/pic/domain/PICCom.F90:40-43, 516-519, 536-539, 556-559, 575-578:
 * ghost = 2*edge - next).  phi is (1:Nz, 1:Nr); Ez (0:Nz, 1:Nr), Er (1:Nz, 0:Nr). */
void ora_set_e(int nz, int nr, double dx, const double *phi, double *Ez, double *Er)
{
    const size_t lz = (size_t)nz + 2;
    double *p = (double *)calloc(lz * (size_t)(nr + 2), sizeof(double));
#define PP(i, j) p[(size_t)(i) + (size_t)(j) * lz]        /* i = 0..nz+1, j = 0..nr+1 */
    for (int j = 1; j <= nr; ++j)
        for (int i = 1; i <= nz; ++i) PP(i, j) = phi[(size_t)(i - 1) + (size_t)(j - 1) * nz];
    for (int j = 1; j <= nr; ++j) {
        PP(0, j) = 2.0 * PP(1, j) - PP(2, j);
        PP(nz + 1, j) = 2.0 * PP(nz, j) - PP(nz - 1, j);
    }
    for (int i = 1; i <= nz; ++i) {
        PP(i, 0) = 2.0 * PP(i, 1) - PP(i, 2);
        PP(i, nr + 1) = 2.0 * PP(i, nr) - PP(i, nr - 1);
    }
    for (int j = 1; j <= nr; ++j)
        for (int i = 0; i <= nz; ++i) {
            double v = PP(i, j) - PP(i + 1, j);
            Ez[(size_t)i + (size_t)(j - 1) * (nz + 1)] = v / dx;
        }
    for (int j = 0; j <= nr; ++j)
        for (int i = 1; i <= nz; ++i) {
            double v = PP(i, j) - PP(i, j + 1);
            Er[(size_t)(i - 1) + (size_t)j * nz] = v / dx;
        }
#undef PP
    free(p);
}

/* The r-z Poisson coefficients, R/src/pic/field/FieldSolver.f90:174-301:
 * SetFieldSolverCoe (with_chi = 1, lines 174-248) or
 * SetFieldSolverCoeWithoutChi (with_chi = 0, lines 251-301), restated loop
 * for loop on a one-rank domain (zstart..zend =
 * 1..Nz, rstart..rend = 1..Nr).  chi is Chi(0:Nz+1, 0:Nr+1); cell_epsilon is
 * (1:Nz-1, 1:Nr-1) and read at the Fortran address the reference computes
 * (EPSR below), so the axis term at i = Nz-1 that names cell (Nz, 1) reads
 * the element stored there, cell (1, 2).  coe = CoeA..CoeE, each Nz*Nr. */
void ora_solver_coe(int nz, int nr, int with_chi, const double *chi, const double *cell_epsilon,
                    const int *point_type, double *coe)
{
    const size_t nn = (size_t)nz * nr;
    double *coeA = coe, *coeB = coe + nn, *coeC = coe + 2 * nn, *coeD = coe + 3 * nn, *coeE = coe + 4 * nn;
#define CHI(i, j) chi[(size_t)(i) + (size_t)(j) * (size_t)(nz + 2)]
#define EPSR(i, j) cell_epsilon[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)(nz - 1)]
#define AT(a, i, j) a[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)nz]
    for (size_t k = 0; k < 5 * nn; ++k) coe[k] = 0.0;
    for (int j = 1; j <= nr; ++j)
        for (int i = 1; i <= nz; ++i) {
            const int point_type_one = AT(point_type, i, j);
            if (point_type_one > 0) {
                AT(coeB, i, j) = 1.0;
            } else if (nr == j || 1 == i || nz == i) {
                /* boundary: not used */
            } else if (1 == j) {
                if (with_chi) {
                    AT(coeA, i, j) = (1.0 + 0.5 * (CHI(i, j) + CHI(i - 1, j))) * 0.5 * (EPSR(i, j) + EPSR(i - 1, j));
                    AT(coeC, i, j) = (1.0 + 0.5 * (CHI(i, j) + CHI(i + 1, j))) * 0.5 * (EPSR(i, j) + EPSR(i + 1, j));
                    AT(coeE, i, j) =
                        4.0 * (1.0 + 0.5 * (CHI(i, j) + CHI(i, j + 1))) * 0.5 * (EPSR(i, j) + EPSR(i, j + 1));
                } else {
                    AT(coeA, i, j) = 0.5 * (EPSR(i, j) + EPSR(i - 1, j));
                    AT(coeC, i, j) = 0.5 * (EPSR(i, j) + EPSR(i + 1, j));
                    AT(coeE, i, j) = 4.0 * 0.5 * (EPSR(i, j) + EPSR(i, j + 1));
                }
                AT(coeB, i, j) = -(AT(coeA, i, j) + AT(coeC, i, j) + AT(coeD, i, j) + AT(coeE, i, j));
            } else {
                if (with_chi) {
                    AT(coeA, i, j) =
                        (1.0 + 0.5 * (CHI(i, j) + CHI(i - 1, j))) * 0.5 * (EPSR(i - 1, j - 1) + EPSR(i - 1, j));
                    AT(coeC, i, j) = (1.0 + 0.5 * (CHI(i, j) + CHI(i + 1, j))) * 0.5 * (EPSR(i, j - 1) + EPSR(i, j));
                    AT(coeD, i, j) = (1.0 + 0.5 * (CHI(i, j) + CHI(i, j - 1))) * 0.5 *
                                     (EPSR(i - 1, j - 1) + EPSR(i, j - 1)) * (j - 1.5) / (j - 1.0);
                    AT(coeE, i, j) = (1.0 + 0.5 * (CHI(i, j) + CHI(i, j + 1))) * 0.5 * (EPSR(i - 1, j) + EPSR(i, j)) *
                                     (j - 0.5) / (j - 1.0);
                } else {
                    AT(coeA, i, j) = 0.5 * (EPSR(i - 1, j - 1) + EPSR(i - 1, j));
                    AT(coeC, i, j) = 0.5 * (EPSR(i, j - 1) + EPSR(i, j));
                    AT(coeD, i, j) = 0.5 * (EPSR(i - 1, j - 1) + EPSR(i, j - 1)) * (j - 1.5) / (j - 1.0);
                    AT(coeE, i, j) = 0.5 * (EPSR(i - 1, j) + EPSR(i, j)) * (j - 0.5) / (j - 1.0);
                }
                AT(coeB, i, j) = -(AT(coeA, i, j) + AT(coeC, i, j) + AT(coeD, i, j) + AT(coeE, i, j));
            }
        }
#undef CHI
#undef EPSR
#undef AT
}

/* SetFieldSolverSource then SetSolverBoundary, FieldSolver.f90:153-171 and
 * 136-150:
 * Source = -1 * dx * dx / Epsilon * (Rho + sigma * inv_vol) (Fortran's
 * unary minus applies to the whole product), then Source = bound where
 * bound_mask /= 0.  sigma, bound and bound_mask may be NULL (zero / no
 * merge).  Every array is Nz*Nr. */
void ora_solver_source(int nz, int nr, double dx, const double *rho, const double *sigma, const double *inv_vol,
                       const double *bound, const int *bound_mask, double *source)
{
    const size_t nn = (size_t)nz * nr;
    for (size_t k = 0; k < nn; ++k) {
        const double sg = sigma ? sigma[k] : 0.0;
        source[k] = -(1.0 * dx * dx / C_EPS0 * (rho[k] + sg * inv_vol[k]));
    }
    if (bound && bound_mask)
        for (size_t k = 0; k < nn; ++k)
            if (bound_mask[k] != 0) source[k] = bound[k];
}

/* WeightingParticleField for one species, R/src/diag/diagParticleField.F90:196-257
 * (node sums) and 300-318 (scaling).  P holds n particles (no deleted
 * slots).  implicit_j: the currents use a copy pushed once more in (Ez, Er)
 * (MoveParticleOneElectrostaticImplicit, ParticleBundle.f90:782-858); else
 * the particle's own Vz and Vr (the reference's explicit branch reads Vr of a
 * temporary it never assigns, diagParticleField.F90:249-252). */
void ora_moments(int nz, int nr, double dx, double dt, double charge, double mass, const double *Ez,
                 const double *Er, int implicit_j, int64_t n, double *const P[8], const double *inv_vol,
                 double *rho, double *energy, double *mp, double *jz, double *jr)
{
    const size_t nn = (size_t)nz * nr;
    const double VFactor = dx / dt;     /* ParticleBundle.f90:346 */
    for (size_t k = 0; k < nn; ++k) rho[k] = energy[k] = mp[k] = jz[k] = jr[k] = 0.0;
#define AT(a, i, j) a[((size_t)(i) - 1) + ((size_t)(j) - 1) * nz]
    for (int64_t p = 0; p < n; ++p) {
        const double Z = P[0][p], R = P[1][p], WQ = P[7][p];
        int NzL = (int)ceil(Z);
        int NzU = NzL + 1;
        double ZL = (double)NzL - Z;
        double ZU = 1.0 - ZL;
        int NrL = (int)ceil(R);
        int NrU = NrL + 1;
        double RL = (double)NrL - R;
        double RU = 1.0 - RL;
        double Energy = P[3][p] * P[3][p] + P[4][p] * P[4][p] + P[2][p] * P[2][p];
        double tz = P[2][p], tr = P[3][p];
        if (implicit_j) {
            double z1 = Z, r1 = R, vz = P[2][p], vr = P[3][p], vt = P[4][p], az = P[5][p], ar = P[6][p];
            ora_push(nz, nr, dx, dt, charge, mass, Ez, Er, 1, &z1, &r1, &vz, &vr, &vt, &az, &ar);
            tz = vz;
            tr = vr;
        }
        /* R == 0 exactly: row NrL = 0 lies outside the arrays and its weight RL is 0 */
        if (NrL >= 1) {
            AT(rho, NzL, NrL) = AT(rho, NzL, NrL) + ZL * RL * WQ;
            AT(energy, NzL, NrL) = AT(energy, NzL, NrL) + ZL * RL * WQ * Energy;
            AT(mp, NzL, NrL) = AT(mp, NzL, NrL) + ZL * RL;
            AT(jz, NzL, NrL) = AT(jz, NzL, NrL) + ZL * RL * WQ * tz;
            AT(jr, NzL, NrL) = AT(jr, NzL, NrL) + ZL * RL * WQ * tr;
        }
        AT(rho, NzL, NrU) = AT(rho, NzL, NrU) + ZL * RU * WQ;
        AT(energy, NzL, NrU) = AT(energy, NzL, NrU) + ZL * RU * WQ * Energy;
        AT(mp, NzL, NrU) = AT(mp, NzL, NrU) + ZL * RU;
        AT(jz, NzL, NrU) = AT(jz, NzL, NrU) + ZL * RU * WQ * tz;
        AT(jr, NzL, NrU) = AT(jr, NzL, NrU) + ZL * RU * WQ * tr;
        if (NrL >= 1) {
            AT(rho, NzU, NrL) = AT(rho, NzU, NrL) + ZU * RL * WQ;
            AT(energy, NzU, NrL) = AT(energy, NzU, NrL) + ZU * RL * WQ * Energy;
            AT(mp, NzU, NrL) = AT(mp, NzU, NrL) + ZU * RL;
            AT(jz, NzU, NrL) = AT(jz, NzU, NrL) + ZU * RL * WQ * tz;
            AT(jr, NzU, NrL) = AT(jr, NzU, NrL) + ZU * RL * WQ * tr;
        }
        AT(rho, NzU, NrU) = AT(rho, NzU, NrU) + ZU * RU * WQ;
        AT(energy, NzU, NrU) = AT(energy, NzU, NrU) + ZU * RU * WQ * Energy;
        AT(mp, NzU, NrU) = AT(mp, NzU, NrU) + ZU * RU;
        AT(jz, NzU, NrU) = AT(jz, NzU, NrU) + ZU * RU * WQ * tz;
        AT(jr, NzU, NrU) = AT(jr, NzU, NrU) + ZU * RU * WQ * tr;
    }
#undef AT
    const double EnergyFactor = 0.5 * mass * VFactor * VFactor;
    for (size_t k = 0; k < nn; ++k) {
        rho[k] = rho[k] * inv_vol[k];
        energy[k] = energy[k] * inv_vol[k];
        energy[k] = energy[k] * EnergyFactor;
        jz[k] = jz[k] * inv_vol[k] * charge * VFactor;
        jr[k] = jr[k] * inv_vol[k] * charge * VFactor;
        if (rho[k] > 0.0) energy[k] = energy[k] / C_JTOEV / rho[k];
    }
}

/* ---------------------------------------------------------------- MCC */
typedef struct { double Z, R, Vz, Vr, Vt, Az, Ar, WQ; } po_t;

typedef struct {
    double miu, mass_ratio, energy, beta;
    double gx, gy, gz, gper, g;
    po_t poi, pot;
    int nnew;
    po_t pon[2];
    int pon_idx[2];
    int pon_twin[2];   /* product k duplicates another particle of its species */
} mcpo_t;

/* VecRanXYZInit, ParticleOne.f90:108-123 */
static void vec_ran_xyz(po_t *p, double V, rng_t *rg)
{
    double R = rnext(rg);
    double CosTheta = 1.0 - 2.0 * R;
    double SinTheta = sqrt(1.0 - CosTheta * CosTheta);
    R = rnext(rg);
    double Phi = 2.0 * C_PI * R;
    p->Vz = V * CosTheta;
    p->Vr = V * SinTheta * cos(Phi);
    p->Vt = V * SinTheta * sin(Phi);
}

/* VelMaxInit, ParticleOne.f90:76-95 */
static void vel_max_init(po_t *p, double Mass, double Temperature, rng_t *rg)
{
    double Beta = 1.0 / (C_KB * Temperature);
    double FuncA = 1.0, FuncB = 0.0, R = 0.0;
    while (FuncA > FuncB) {
        R = rnext(rg);
        FuncA = R * R;
        R = rnext(rg);
        FuncB = -C_EXP * R * log(R);
    }
    double V = sqrt(-3.0 * log(R) / Beta / Mass);
    vec_ran_xyz(p, V, rg);
}

uint32_t ora_vel_maxwell(double mass, double temperature, uint64_t seed, uint32_t slot,
                         uint32_t step, uint32_t draw, uint32_t tag,
                         double *vz, double *vr, double *vt)
{
    rng_t rg = {seed, step, tag, slot, 0, 0, draw};
    po_t p;
    memset(&p, 0, sizeof p);
    vel_max_init(&p, mass, temperature, &rg);
    *vz = p.Vz; *vr = p.Vr; *vt = p.Vt;
    return rg.k;
}

/* Updater, TypeMCC.f90:60-92: the target sample, then the relative terms */
static void mc_relative(mcpo_t *m, double sm, double gm);
static void mc_update(mcpo_t *m, double sm, double gm, double gT, rng_t *rg)
{
    m->pot = m->poi;
    m->miu = sm * gm / (sm + gm);
    m->mass_ratio = gm / (sm + gm);
    vel_max_init(&m->pot, gm, gT, rg);
    mc_relative(m, sm, gm);
}

/* TypeMCC.f90:78-91: relative velocity, CM energy, target Az/Ar */
static void mc_relative(mcpo_t *m, double sm, double gm)
{
    m->gx = m->poi.Vr - m->pot.Vr;
    m->gy = m->poi.Vt - m->pot.Vt;
    m->gz = m->poi.Vz - m->pot.Vz;
    double Gtemp = m->gy * m->gy + m->gz * m->gz;
    m->gper = sqrt(Gtemp);
    double G2 = Gtemp + m->gx * m->gx;
    m->g = sqrt(G2);
    m->energy = 0.5 * m->miu * G2 / C_JTOEV;
    m->pot.Ar = m->pot.Ar * sm / gm;
    m->pot.Az = m->pot.Az * sm / gm;
}

/* VelocityUpdater, TypeMCC.f90:94-112 */
static void mc_vel_update(mcpo_t *m, double VF)
{
    m->gx = VF * m->gx; m->gy = VF * m->gy; m->gz = VF * m->gz;
    m->gper = VF * m->gper; m->g = VF * m->g;
    m->poi.Vr = VF * m->poi.Vr; m->poi.Vt = VF * m->poi.Vt; m->poi.Vz = VF * m->poi.Vz;
}

/* MCCChi.f90:5-25 */
static double cos_isotropic(double E, rng_t *rg) { (void)E; double R = rnext(rg); return 1.0 - 2.0 * R; }
static double cos_vahedi(double E, rng_t *rg)
{
    if (E < C_MINREAL) return 1.0;
    double R = rnext(rg);
    return (E + 2.0 - 2.0 * pow(E + 1.0, R)) / E;
}
/* MCCChi.f90:49-59 */
static double create_energy_vahedi(double E, rng_t *rg)
{
    double R = rnext(rg);
    double v = 10.0 * tan(R * atan(E) / 20.0);
    if (fabs(v) < C_MINREAL) v = C_MINREAL;
    return v;
}

/* PostCollisionVelocity, TypeMCC.f90:114-146 */
static void post_collision(mcpo_t *m, int vahedi, rng_t *rg)
{
    if (fabs(m->gper) < C_MINREAL) {
        vec_ran_xyz(&m->poi, m->g, rg);
    } else {
        double CosTheta = vahedi ? cos_vahedi(m->energy, rg) : cos_isotropic(m->energy, rg);
        double FcosTheta = 1.0 - CosTheta;
        double SinTheta = sqrt(1.0 - CosTheta * CosTheta);
        double R = rnext(rg);
        double APhi = 2.0 * C_PI * R;
        double CosPhi = cos(APhi), SinPhi = sin(APhi);
        double gx = m->gx, gy = m->gy, gz = m->gz, g = m->g, gper = m->gper;
        double hx = gper * CosPhi;
        double hy = -(gx * gy * CosPhi + g * gz * SinPhi) / gper;
        double hz = -(gx * gz * CosPhi - g * gy * SinPhi) / gper;
        m->poi.Vr = m->poi.Vr - m->mass_ratio * (gx * FcosTheta + hx * SinTheta);
        m->poi.Vt = m->poi.Vt - m->mass_ratio * (gy * FcosTheta + hy * SinTheta);
        m->poi.Vz = m->poi.Vz - m->mass_ratio * (gz * FcosTheta + hz * SinTheta);
    }
}

/* SelectProbility, TypeMCC.f90:173-216, with its uniform R passed in: the
 * stream's draw 1, drawn before Updater's target sampling (draws 2 on). */
static int select_probability(const mcpo_t *m, const ora_mcc_table *t, double R)
{
    double tmp[256];
    const int Nr = t->nreaction, Ns = t->nsigma;
    double E = m->energy;
    double lx = (log10(E) - t->emin) / t->einterval;
    int N = (E > 0.0) ? (int)ceil(lx) : 0;
    if (N < 1) {
        for (int j = 0; j < Nr; ++j) tmp[j] = t->prob[j];
    } else if (N < Ns) {
        double S1 = (double)N - (log10(E) - t->emin) / t->einterval;
        double S2 = 1.0 - S1;
        const double *a = t->prob + (size_t)(N - 1) * Nr, *b = t->prob + (size_t)N * Nr;
        for (int j = 0; j < Nr; ++j) tmp[j] = a[j] * S1 + b[j] * S2;
    } else {
        const double *a = t->prob + (size_t)(Ns - 1) * Nr;
        for (int j = 0; j < Nr; ++j) tmp[j] = a[j] * E / t->emax;
    }
    for (int j = 0; j < Nr; ++j)
        if (E < t->threshold[j]) tmp[j] = 0.0;
    for (int j = 0; j < Nr; ++j)
        if (R < tmp[j]) return j + 1;
    return 0;
}

/* SelectProbilityReactive, TypeMCC.f90:218-236: R (beta's draw) is the
 * stream's draw 1 as above; the non-reactive choice takes the next draw. */
static int select_probability_reactive(mcpo_t *m, const ora_mcc_table *t, double R, rng_t *rg)
{
    m->beta = t->beta_max * sqrt(R);
    if (m->beta <= 1.0) {
        if (m->energy <= t->threshold[1]) return 1;
        return select_probability(m, t, rnext(rg));
    }
    return 1;
}

/* SelectCollisionElectron, MCCElectron.F90:13-147 */
static void collide_electron(mcpo_t *m, int species, int rtype, int resultant, double th, rng_t *rg)
{
    switch (rtype) {
    case 11:
        post_collision(m, 1, rg);
        break;
    case 12: {
        double VF = sqrt(1.0 - th / m->energy);
        m->energy = m->energy - th;
        mc_vel_update(m, VF);
        post_collision(m, 1, rg);
    } break;
    case 13: {
        po_t solid = m->poi;
        double E0 = m->energy;
        double EnergyTemp = m->energy - th;
        double EnergyOld = create_energy_vahedi(EnergyTemp, rg);
        double VF = sqrt(EnergyOld / m->energy);
        mc_vel_update(m, VF);
        post_collision(m, 1, rg);
        double EnergyNew = EnergyTemp - EnergyOld;
        double VF2 = sqrt(EnergyNew / E0);
        m->nnew = 2;
        m->pon[0] = solid;
        m->pon[0].Vz = m->pon[0].Vz * VF2;
        m->pon[0].Vr = m->pon[0].Vr * VF2;
        m->pon[0].Vt = m->pon[0].Vt * VF2;
        m->pon_idx[0] = species;
        m->pon[1] = m->pot;
        m->pon_idx[1] = resultant;
    } break;
    case 14:
        m->nnew = 1;
        m->pon[0] = m->pot;
        m->pon_idx[0] = resultant;
        break;
    case 15: {
        double VF = sqrt(1.0 - th / m->energy);
        m->energy = m->energy - th;
        mc_vel_update(m, VF);
        post_collision(m, 1, rg);
        m->nnew = 2;
        m->pon[0] = m->pot;
        m->pon[1] = m->pot;
        m->pon_idx[0] = resultant % 10;
        m->pon_idx[1] = resultant / 10;
        m->pon_twin[1] = (resultant % 10 == resultant / 10);
    } break;
    default:
        break;
    }
}

/* PostVelocityIonReactive, MCCIon.F90:54-74 (Vy takes Vz, as in the source) */
static void ion_reactive(mcpo_t *m, double th)
{
    double VF = sqrt(1.0 - th / m->energy);
    m->energy = m->energy - th;
    mc_vel_update(m, VF);
    double A = m->mass_ratio, B = 1.0 - A;
    m->poi.Vr = B * m->poi.Vr + A * m->pot.Vr + B * m->gx;
    m->poi.Vt = B * m->poi.Vz + A * m->pot.Vt + B * m->gy;
    m->poi.Vz = B * m->poi.Vz + A * m->pot.Vz + B * m->gz;
}

/* SelectCollisionIon, MCCIon.F90:6-51 */
static void collide_ion(mcpo_t *m, int species, int rtype, int resultant, double th, rng_t *rg)
{
    switch (rtype) {
    case 101: post_collision(m, 0, rg); break;
    case 102:
    case 103: m->poi = m->pot; break;
    case 100: break;
    case 111: ion_reactive(m, th); break;
    case 112:
        ion_reactive(m, th);
        m->nnew = 1; m->pon[0] = m->pot; m->pon_idx[0] = resultant;
        break;
    case 113:
        ion_reactive(m, th);
        m->nnew = 1; m->pon[0] = m->poi; m->pon_idx[0] = resultant;
        m->pon_twin[0] = (resultant == species);
        break;
    default: break;
    }
}

/* Test hook: reaction `rtype` of one particle against a given target.
 * poi: Z R Vz Vr Vt Az Ar WQ with velocities in m/s (after VelRes(VFactor));
 * pot_v: the target's Vz Vr Vt in m/s, in place of the Maxwellian sample
 * Updater draws (TypeMCC.f90:60-77); the relative terms follow
 * (TypeMCC.f90:78-91), then SelectCollisionElectron (model 1) or
 * SelectCollisionIon (models 2, 3).  Draws the collision needs come from
 * the stream (seed, step 0, tag 0, c0, 0, 0).  out: the particle afterwards;
 * energy_out: MCPO%Energy afterwards; prod/prod_species: the products.
 * Returns the number of products. */
int ora_collide_given(int model, int species, int rtype, int resultant, double threshold, double species_mass,
                      double gas_mass, const double poi[8], const double pot_v[3], uint64_t seed, uint32_t c0,
                      double out[8], double *energy_out, double prod[16], int prod_species[2])
{
    mcpo_t m;
    memset(&m, 0, sizeof m);
    m.poi.Z = poi[0]; m.poi.R = poi[1]; m.poi.Vz = poi[2]; m.poi.Vr = poi[3];
    m.poi.Vt = poi[4]; m.poi.Az = poi[5]; m.poi.Ar = poi[6]; m.poi.WQ = poi[7];
    m.pot = m.poi;
    m.miu = species_mass * gas_mass / (species_mass + gas_mass);
    m.mass_ratio = gas_mass / (species_mass + gas_mass);
    m.pot.Vz = pot_v[0]; m.pot.Vr = pot_v[1]; m.pot.Vt = pot_v[2];
    mc_relative(&m, species_mass, gas_mass);
    rng_t rg = {seed, 0, 0, c0, 0, 0, 0};
    if (model == 1) collide_electron(&m, species, rtype, resultant, threshold, &rg);
    else collide_ion(&m, species, rtype, resultant, threshold, &rg);
    const po_t *q = &m.poi;
    out[0] = q->Z; out[1] = q->R; out[2] = q->Vz; out[3] = q->Vr;
    out[4] = q->Vt; out[5] = q->Az; out[6] = q->Ar; out[7] = q->WQ;
    *energy_out = m.energy;
    for (int k = 0; k < m.nnew && k < 2; ++k) {
        const po_t *p = &m.pon[k];
        double *o = prod + 8 * k;
        o[0] = p->Z; o[1] = p->R; o[2] = p->Vz; o[3] = p->Vr; o[4] = p->Vt; o[5] = p->Az; o[6] = p->Ar; o[7] = p->WQ;
        prod_species[k] = m.pon_idx[k];
    }
    return m.nnew;
}

int ora_mcc(const ora_mcc_setup *st, uint64_t seed, uint32_t step,
            int64_t *n, const int64_t *cap, double *(*P)[8], int64_t *counts)
{
    const double VFactor = st->dx / st->dt;
    const double invVF = 1.0 / VFactor;
    int64_t nst = 0, nstcap = 1024;
    po_t *stage = (po_t *)malloc(sizeof(po_t) * nstcap);
    int *stage_idx = (int *)malloc(sizeof(int) * nstcap);
    int *stage_twin = (int *)malloc(sizeof(int) * nstcap);
    int rc = 0;
    /* every particle's stream is keyed by its state at the start of MCC */
    uint32_t *keys[64] = {0};
    for (int s = 0; s < st->ns && s < 64; ++s) {
        keys[s] = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (size_t)(n[s] > 0 ? n[s] : 1));
        for (int64_t i = 0; i < n[s]; ++i)
            ora_state_key(P[s][0][i], P[s][1][i], P[s][2][i], P[s][3][i], P[s][4][i], keys[s] + 3 * i);
    }
    for (int g = 0; g < st->ng; ++g)
        for (int s = 0; s < st->ns; ++s) {
            const ora_mcc_table *t = &st->tab[s * st->ng + g];
            if (t->model == 0 || n[s] <= 0) continue;
            const int64_t np = n[s];
            double *const *A = P[s];
            for (int64_t i = 0; i < np; ++i) {
                const uint32_t *kk = keys[s] + 3 * i;
                rng_t rg = {seed, step, (uint32_t)s | ((uint32_t)g << 8), kk[0], kk[1], kk[2], 0};
                double R = rnext(&rg);
                if (!(R < t->collision_ratio)) continue;
                mcpo_t m;
                memset(&m, 0, sizeof m);
                m.poi.Z = A[0][i]; m.poi.R = A[1][i]; m.poi.Vz = A[2][i]; m.poi.Vr = A[3][i];
                m.poi.Vt = A[4][i]; m.poi.Az = A[5][i]; m.poi.Ar = A[6][i]; m.poi.WQ = A[7][i];
                m.poi.Vz = m.poi.Vz * VFactor; m.poi.Vr = m.poi.Vr * VFactor; m.poi.Vt = m.poi.Vt * VFactor;
                /* the reaction choice's uniform is draw 1, the target's draws follow */
                const double u = rnext(&rg);
                mc_update(&m, st->mass[s], t->gas_mass, t->gas_temperature, &rg);
                int idx = 0;
                if (t->model == 1 || t->model == 2) idx = select_probability(&m, t, u);
                else if (t->model == 3) idx = select_probability_reactive(&m, t, u, &rg);
                if (idx > 0) {
                    int ty = t->rtype[idx - 1];
                    if (t->model == 1)
                        collide_electron(&m, s, ty, t->resultant[idx - 1], t->threshold[idx - 1], &rg);
                    else
                        collide_ion(&m, s, ty, t->resultant[idx - 1], t->threshold[idx - 1], &rg);
                }
                if (counts) counts[((int64_t)s * st->ng + g) * 128 + (idx & 127)] += 1;
                m.poi.Vz = m.poi.Vz * invVF; m.poi.Vr = m.poi.Vr * invVF; m.poi.Vt = m.poi.Vt * invVF;
                A[0][i] = m.poi.Z; A[1][i] = m.poi.R; A[2][i] = m.poi.Vz; A[3][i] = m.poi.Vr;
                A[4][i] = m.poi.Vt; A[5][i] = m.poi.Az; A[6][i] = m.poi.Ar; A[7][i] = m.poi.WQ;
                for (int k = 0; k < m.nnew; ++k) {
                    if (nst == nstcap) {
                        nstcap *= 2;
                        stage = (po_t *)realloc(stage, sizeof(po_t) * nstcap);
                        stage_idx = (int *)realloc(stage_idx, sizeof(int) * nstcap);
                        stage_twin = (int *)realloc(stage_twin, sizeof(int) * nstcap);
                    }
                    stage[nst] = m.pon[k];
                    stage_idx[nst] = m.pon_idx[k];
                    stage_twin[nst] = m.pon_twin[k];
                    ++nst;
                }
            }
        }
    /* MCCAddNewParticle, MCCPublic.F90:209-226 */
    for (int64_t k = 0; k < nst; ++k) {
        int d = stage_idx[k];
        if (d < 0 || d >= st->ns || n[d] >= cap[d]) { rc = -1; continue; }
        po_t p = stage[k];
        p.Vz = p.Vz * invVF; p.Vr = p.Vr * invVF; p.Vt = p.Vt * invVF;
        /* a twin of another particle of its species made by the same
         * collision: its Vt pattern's low bit is flipped, so that the two
         * key different streams afterwards (not in the reference, whose
         * global random stream separates such twins by itself) */
        if (stage_twin[k]) p.Vt = flip_low_bit(p.Vt);
        int64_t j = n[d]++;
        P[d][0][j] = p.Z; P[d][1][j] = p.R; P[d][2][j] = p.Vz; P[d][3][j] = p.Vr;
        P[d][4][j] = p.Vt; P[d][5][j] = p.Az; P[d][6][j] = p.Ar; P[d][7][j] = p.WQ;
    }
    free(stage);
    free(stage_idx);
    free(stage_twin);
    for (int s = 0; s < st->ns && s < 64; ++s) free(keys[s]);
    return rc;
}

/* ------------------------------------------------ cross-section tables */
int ora_locate(const double *xx, int n, double x)
{
    /* xx is 1-based in the reference; here xx[0] is xx(1) */
    int jl = 0, ju = n + 1;
    while (ju - jl > 1) {
        int jm = (ju + jl) / 2;
        if ((xx[n - 1] >= xx[0]) == (x >= xx[jm - 1])) jl = jm;
        else ju = jm;
    }
    if (x == xx[0]) return 1;
    if (x == xx[n - 1]) return n - 1;
    return jl;
}

#define NSIGMA_RAW_MAX 500

void ora_sigma_normalize(int nreaction, int nsigma, double emin, double einterval,
                         const double *threshold, const double *raw_e, const double *raw_s,
                         double *sigma_out)
{
    double xx[NSIGMA_RAW_MAX + 2];
    for (int j = 0; j < nreaction; ++j) {
        const double *e = raw_e + (size_t)j * NSIGMA_RAW_MAX;
        const double *s = raw_s + (size_t)j * NSIGMA_RAW_MAX;
        int nraw = 1;   /* Maxloc: first location of the maximum, 1-based */
        for (int i = 2; i <= NSIGMA_RAW_MAX; ++i)
            if (e[i - 1] > e[nraw - 1]) nraw = i;
        double MaxEnergy = e[nraw - 1];
        double Alfa = s[nraw - 1] * MaxEnergy;
        for (int i = 0; i < NSIGMA_RAW_MAX; ++i) xx[i] = e[i];
        xx[NSIGMA_RAW_MAX] = 0.0; xx[NSIGMA_RAW_MAX + 1] = 0.0;
        for (int i = 1; i <= nraw; ++i) {
            if (xx[i - 1] < pow(10.0, emin)) xx[i - 1] = emin;
            else xx[i - 1] = (log10(xx[i - 1]) - emin) / einterval;
        }
        double Th = threshold[j];
        for (int i = 1; i <= nsigma; ++i) {
            double NEnergy = (double)(i - 1);
            double Energy = pow(10.0, emin + (double)(i - 1) * einterval);
            double v;
            if (Energy <= Th) v = 0.0;
            else if (Energy >= MaxEnergy) v = Alfa / Energy;
            else {
                int k = ora_locate(xx, nraw, NEnergy);
                if (k < 1) k = 1;
                double LE = xx[k - 1], UE = xx[k];
                double S1 = ((double)i - LE) / (UE - LE);
                double S2 = 1.0 - S1;
                double sk1 = (k < NSIGMA_RAW_MAX) ? s[k] : 0.0;
                v = S2 * s[k - 1] + S1 * sk1;
            }
            sigma_out[(size_t)(i - 1) * nreaction + j] = v;
        }
    }
}

void ora_sigma_normalize_reactive(int nreaction, int nsigma, double emin, double einterval,
                                  const double *threshold, double natom_species, double natom_gas,
                                  double *sigma_out)
{
    double NComplex = natom_species + natom_gas;
    double S = (3.0 * NComplex - 6.0) / 2.0 - 1.0;
    double PTemp[256];
    for (int i = 1; i <= nsigma; ++i) {
        double Energy = pow(10.0, emin + (double)(i - 1) * einterval);
        double PSum = 0.0;
        for (int j = 0; j < nreaction; ++j) {
            PTemp[j] = (Energy > threshold[j]) ? pow(Energy - threshold[j], S) : 0.0;
        }
        for (int j = 0; j < nreaction; ++j) PSum = PSum + PTemp[j];
        for (int j = 0; j < nreaction; ++j) sigma_out[(size_t)(i - 1) * nreaction + j] = PTemp[j] / PSum;
    }
}

double ora_probability_nonreactive(int nreaction, int nsigma, double emin, double einterval,
                                   double species_mass, double gas_mass, double gas_density,
                                   const double *sigma, double *prob)
{
    double Miu = species_mass * gas_mass / (species_mass + gas_mass);
    for (int i = 1; i <= nsigma; ++i) {
        double Energy = pow(10.0, emin + (double)(i - 1) * einterval);
        double V = sqrt(2.0 * Energy * C_JTOEV / Miu);
        for (int j = 0; j < nreaction; ++j) {
            size_t id = (size_t)(i - 1) * nreaction + j;
            prob[id] = sigma[id] * V * gas_density;
        }
    }
    for (int i = 1; i <= nsigma; ++i)
        for (int j = 1; j < nreaction; ++j) {
            size_t id = (size_t)(i - 1) * nreaction + j;
            prob[id] = prob[id] + prob[id - 1];
        }
    double Max = prob[0];
    size_t nt = (size_t)nsigma * nreaction;
    for (size_t k = 1; k < nt; ++k) if (prob[k] > Max) Max = prob[k];
    for (size_t k = 0; k < nt; ++k) prob[k] = prob[k] / Max;
    return Max;
}

double ora_probability_reactive(int nreaction, int nsigma, double species_mass, double gas_mass,
                                double gas_density, double beta_max, double polarizability,
                                double dt, const double *sigma, double *prob)
{
    size_t nt = (size_t)nsigma * nreaction;
    for (size_t k = 0; k < nt; ++k) prob[k] = sigma[k];
    double Miu = species_mass * gas_mass / (species_mass + gas_mass);
    double ratio = sqrt(C_PI * polarizability * C_E * C_E / (C_EPS0 * Miu)) * beta_max * beta_max * gas_density * dt;
    for (int i = 1; i <= nsigma; ++i)
        for (int j = 1; j < nreaction; ++j) {
            size_t id = (size_t)(i - 1) * nreaction + j;
            prob[id] = prob[id] + prob[id - 1];
        }
    return ratio;
}

/* ProbilityHardShpere, MCCInit.F90:172-270: Probility(j,i) = base_j *
 * SigmaMax * V(E_i) * n with base (0.5, 1.0) for two reactions (101, 103)
 * or (1.0) for one (101), SigmaMax = 2 pi (r_g+r_s)/2 (r_g+r_s)/2, then
 * normalised by the table's maximum, which is returned. */
double ora_probability_hard_sphere(int nreaction, int nsigma, double emin, double einterval, double species_mass,
                                   double gas_mass, double gas_density, double r_species, double r_gas,
                                   double *prob)
{
    double SigmaMax = 2.0 * C_PI * 0.5 * (r_gas + r_species) * 0.5 * (r_gas + r_species);
    double Miu = species_mass * gas_mass / (species_mass + gas_mass);
    for (int i = 1; i <= nsigma; ++i) {
        prob[(size_t)(i - 1) * nreaction] = (nreaction == 2) ? 0.5 : 1.0;
        if (nreaction == 2) prob[(size_t)(i - 1) * nreaction + 1] = 1.0;
    }
    for (int i = 1; i <= nsigma; ++i) {
        double Energy = pow(10.0, emin + (double)(i - 1) * einterval);
        double V = sqrt(2.0 * Energy * C_JTOEV / Miu);
        for (int j = 0; j < nreaction; ++j) {
            size_t id = (size_t)(i - 1) * nreaction + j;
            prob[id] = prob[id] * SigmaMax * V * gas_density;
        }
    }
    size_t nt = (size_t)nsigma * nreaction;
    double Max = prob[0];
    for (size_t k = 1; k < nt; ++k) if (prob[k] > Max) Max = prob[k];
    for (size_t k = 0; k < nt; ++k) prob[k] = prob[k] / Max;
    return Max;
}

/* ------------------------------------------------ one step, in place ----
 * The reference's OneStep particle sequence (R/src/onestep.F90:141-177):
 * for every species MoveESIm then MaterialAborptionParticle, then P2CES raw
 * deposits, then MCC.  P[s] arrays have capacity cap[s]; n[s] is updated.
 * rho_raw: ns*nz*nr (accumulated, caller clears). Returns ora_mcc's code. */
int ora_step(const ora_geom *g, const ora_mcc_setup *st, const double *Ez, const double *Er,
             uint64_t seed, uint32_t step, int64_t *n, const int64_t *cap, double *(*P)[8],
             double *rho_raw, double *metal_charge, double *sigma, int64_t *counts)
{
    const size_t nn = (size_t)g->nz * g->nr;
    for (int s = 0; s < st->ns; ++s) {
        double **A = P[s];
        ora_push(g->nz, g->nr, st->dx, st->dt, st->charge[s], st->mass[s], Ez, Er, n[s],
                 A[0], A[1], A[2], A[3], A[4], A[5], A[6]);
        ora_absorb(g, st->charge[s], &n[s], A, metal_charge, sigma);
    }
    for (int s = 0; s < st->ns; ++s)
        ora_deposit_raw(g->nz, g->nr, n[s], P[s][0], P[s][1], P[s][7], rho_raw + (size_t)s * nn);
    return ora_mcc(st, seed, step, n, cap, P, counts);
}

/* Exposes the scattering-angle samplers (IsotropicCosKai / CosKaiVahedi,
 * R/src/mc/MCCChi.f90:5-25) and the secondary-energy sampler
 * (CreatEnergyVehadi, MCCChi.f90:49-59) for distribution tests. */
double ora_cos_chi(int vahedi, double energy, uint64_t seed, uint32_t slot, uint32_t draw)
{
    rng_t rg = {seed, 0, 0, slot, 0, 0, draw};
    return vahedi ? cos_vahedi(energy, &rg) : cos_isotropic(energy, &rg);
}

double ora_secondary_energy(double energy, uint64_t seed, uint32_t slot, uint32_t draw)
{
    rng_t rg = {seed, 0, 0, slot, 0, 0, draw};
    return create_energy_vahedi(energy, &rg);
}
