/*
 * CPU oracle for the iPM2D particle hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement, statement by statement, of the reference routines
 * listed below.  It exists to check the HIP product path; nothing in the
 * product links, loads or calls it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may use it.
 *
 * Array conventions follow the reference's Fortran column-major layouts with
 * z fastest (A(iz, ir)):
 *   Ez(0:Nz, 1:Nr)        -> Ez[iz + (ir-1)*(Nz+1)]
 *   Er(1:Nz, 0:Nr)        -> Er[(iz-1) + ir*Nz]
 *   Rho(1:Nz, 1:Nr)       -> Rho[(iz-1) + (ir-1)*Nz]
 *   cell_material(1:Nz-1, 1:Nr-1), hedge_type(1:Nz, 1:Nr-1),
 *   vedge_type(1:Nz-1, 1:Nr)
 * Particles are structure-of-arrays of the 8 ParticleOne fields
 * (R/src/pic/particle/ParticleOne.f90:8): Z R Vz Vr Vt Az Ar WQ, positions in
 * grid units and velocities in grid units per step.
 */
#ifndef IPM_ORACLE_H
#define IPM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG: Philox4x32-10 (Salmon et al., SC'11) ---------- */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* one uniform double in (0,1): draw `draw` of the stream (c0, c1, c2) */
double ora_uniform(uint64_t seed, uint32_t step, uint32_t tag, uint32_t c0, uint32_t c1, uint32_t c2,
                   uint32_t draw);
/* the stream prefix of a particle: 96 bits folded from (Z, R, Vz, Vr, Vt) */
void ora_state_key(double Z, double R, double Vz, double Vr, double Vt, uint32_t out[3]);

/* ---- grid / geometry -------------------------------------------------- */
typedef struct {
    int nz, nr;              /* node counts (GlobalShape) */
    double dx, dt;
    const int *cell_material;  /* (nz-1)*(nr-1) */
    const int *hedge_type;     /* nz*(nr-1) */
    const int *vedge_type;     /* (nz-1)*nr */
    const int *material_type;  /* material_type[m + ndiel], m in [-ndiel, nmetal] */
    int ndiel, nmetal;
} ora_geom;

/* node weighting volumes, R/src/base/init.F90:402-424 */
void ora_node_volumes(int nz, int nr, double dx, const double *factor,
                      double *vol, double *inv_vol);

/* MoveESIm, R/src/pic/particle/ParticleBundle.f90:671-748 */
void ora_push(int nz, int nr, double dx, double dt, double charge, double mass,
              const double *Ez, const double *Er, int64_t n,
              double *Z, double *R, double *Vz, double *Vr, double *Vt,
              double *Az, double *Ar);

/* MaterialAborptionParticle, R/src/pic/particleboundary/ParticleMaterial.F90:10-90,
 * with DelOne's swap-with-last deletion (ParticleBundle.f90:556-576).
 * *n is updated; metal_charge[m-1] and sigma (nz*nr) are accumulated (+=). */
void ora_absorb(const ora_geom *g, double charge, int64_t *n, double *P[8],
                double *metal_charge, double *sigma);

/* P2CES raw accumulation, R/src/pic/field/FieldOne.f90:151-166 (rho_raw += ...) */
void ora_deposit_raw(int nz, int nr, int64_t n, const double *Z, const double *R,
                     const double *WQ, double *rho_raw);
/* P2CES epilogue (FieldOne.f90:168-175): rho = raw*inv_vol*q, chi = rho*ChiFactor */
void ora_deposit_finish(int nz, int nr, double dt, double charge, double mass,
                        const double *rho_raw, const double *inv_vol,
                        double *rho, double *chi);

/* SetE after linear ghost extrapolation (FieldSource.f90:95-120,
 * PICCom.F90:516-578): phi (1:Nz,1:Nr) -> Ez (0:Nz,1:Nr), Er (1:Nz,0:Nr). */
void ora_set_e(int nz, int nr, double dx, const double *phi, double *Ez, double *Er);

/* The host Poisson system, R/src/pic/field/FieldSolver.f90:136-301:
 * SetFieldSolverCoe (with_chi = 1) / SetFieldSolverCoeWithoutChi (0) into
 * coe = CoeA..CoeE (5*Nz*Nr; chi is Chi(0:Nz+1, 0:Nr+1), cell_epsilon
 * (Nz-1)*(Nr-1), point_type Nz*Nr), and SetFieldSolverSource +
 * SetSolverBoundary into source (Nz*Nr; sigma, bound, bound_mask may be NULL). */
void ora_solver_coe(int nz, int nr, int with_chi, const double *chi, const double *cell_epsilon,
                    const int *point_type, double *coe);
void ora_solver_source(int nz, int nr, double dx, const double *rho, const double *sigma, const double *inv_vol,
                       const double *bound, const int *bound_mask, double *source);

/* Recombination (Recombination.F90:164-273): cell densities of one bundle,
 * then the whole call on ns bundles (P: 8 arrays per species). */
void ora_recomb_density(int nz, int nr, double dx, double dt, double mass, int want_t, int64_t n,
                        double *const P[8], double *dens, double *inv_t);
int ora_recombine(int nz, int nr, double dx, double dt, int ns, int64_t *n, double *const *P, const double *mass,
                  int nreaction, const int *reactions, const double *rate, double time_step, uint64_t seed,
                  uint32_t step, int64_t *deleted);

/* split_particle_weight + merge_particle_random (Rezoning.F90:53-92) on one
 * bundle; particle arrays hold cap slots.  -1 when copies exceed cap. */
int ora_split_merge(int64_t *n, double *P[8], int64_t cap, int mode, double weight, double alpha,
                    double grid_r, double merge_limit, uint64_t seed, uint32_t step, int s, int64_t stats[3]);

/* WeightingParticleField for one species (diagParticleField.F90:196-318):
 * rho, energy (mean eV where rho > 0), mp, jz, jr, each nz*nr. */
void ora_moments(int nz, int nr, double dx, double dt, double charge, double mass, const double *Ez,
                 const double *Er, int implicit_j, int64_t n, double *const P[8], const double *inv_vol,
                 double *rho, double *energy, double *mp, double *jz, double *jr);

/* ---- MCC ---------------------------------------------------------------- */
typedef struct {
    int model;            /* 0 none, 1 electron, 2 non-reactive ion, 3 reactive ion */
    int nreaction, nsigma;
    double emin, einterval, emax;   /* emax = 10**EnergyMax */
    double collision_ratio;
    const int *rtype;       /* ReactionType per reaction */
    const int *resultant;   /* Resultant species index per reaction */
    const double *threshold;
    const double *prob;     /* Probility(j,i) -> prob[(i-1)*nreaction + (j-1)] */
    double gas_mass, gas_temperature, beta_max;
} ora_mcc_table;

typedef struct {
    int ns;       /* number of species (0 = electrons) */
    int ng;       /* number of gases */
    double dx, dt;
    const double *charge, *mass;     /* per species */
    const ora_mcc_table *tab;        /* tab[s*ng + g] */
} ora_mcc_setup;

/* MCC, R/src/mc/MCCPublic.F90:38-226 and callees.  Draws come from
 * ora_uniform(seed, step, species|gas<<8, key(state at MCC start), k), k
 * counting draws per particle and gas: k = 0 the selection, k = 1 the
 * uniform of the reaction choice (SelectProbility's R, or the Langevin
 * beta's), k >= 2 Updater's target sampling and then the collision's own.
 * Species arrays P[s][0..7] have capacity cap[s]; n[s] updated with products
 * appended at the end (in reference order).  counts[(s*ng + g)*128 + j]
 * (if non-NULL) tallies events per species, gas and reaction index j
 * (j = 0: null event).  Returns 0, or -1 on capacity overflow. */
int ora_mcc(const ora_mcc_setup *st, uint64_t seed, uint32_t step,
            int64_t *n, const int64_t *cap, double *(*P)[8], int64_t *counts);

/* Test hook: reaction rtype (MCCElectron.F90:13-147 for model 1, MCCIon.F90:6-74
 * otherwise) of one particle poi (Z R Vz Vr Vt Az Ar WQ, velocities in m/s)
 * against the given target velocity pot_v (Vz Vr Vt, m/s) in place of the
 * Maxwellian sample.  Returns the number of products (prod: 8 per product). */
int ora_collide_given(int model, int species, int rtype, int resultant, double threshold, double species_mass,
                      double gas_mass, const double poi[8], const double pot_v[3], uint64_t seed, uint32_t c0,
                      double out[8], double *energy_out, double prod[16], int prod_species[2]);

/* Maxwellian speed + isotropic direction, ParticleOne.f90:76-123, with draws
 * from a per-particle stream; returns the next draw index. */
uint32_t ora_vel_maxwell(double mass, double temperature, uint64_t seed, uint32_t slot,
                         uint32_t step, uint32_t draw, uint32_t tag,
                         double *vz, double *vr, double *vt);

/* OneStep's particle sequence in place: push + absorb per species, raw
 * deposits (rho_raw: ns*nz*nr, accumulated), then MCC. */
int ora_step(const ora_geom *g, const ora_mcc_setup *st, const double *Ez, const double *Er,
             uint64_t seed, uint32_t step, int64_t *n, const int64_t *cap, double *(*P)[8],
             double *rho_raw, double *metal_charge, double *sigma, int64_t *counts);

/* Scattering-angle and secondary-energy samplers, one draw each. */
double ora_cos_chi(int vahedi, double energy, uint64_t seed, uint32_t slot, uint32_t draw);
double ora_secondary_energy(double energy, uint64_t seed, uint32_t slot, uint32_t draw);

/* ---- cross-section tables (R/src/mc/MCCSigma.f90, MCCInit.F90) --------- */
/* locate, R/src/base/numerical/Numrical.F90:112-145 (1-based result) */
int ora_locate(const double *xx, int n, double x);

/* SigmaNormalizedUpdate (MCCSigma.f90:131-190) for Model 1/2.
 * raw_e/raw_s: NSigmaRawMax(=500) entries per reaction, zero padded.
 * sigma_out: Sigma(j,i) -> sigma_out[(i-1)*nreaction + (j-1)] */
void ora_sigma_normalize(int nreaction, int nsigma, double emin, double einterval,
                         const double *threshold, const double *raw_e, const double *raw_s,
                         double *sigma_out);
/* SigmaNormalizedUpdateReactive (MCCSigma.f90:192-220) */
void ora_sigma_normalize_reactive(int nreaction, int nsigma, double emin, double einterval,
                                  const double *threshold, double natom_species, double natom_gas,
                                  double *sigma_out);
/* ProbilityNonReactive (MCCInit.F90:125-170); returns SigmaMax */
double ora_probability_nonreactive(int nreaction, int nsigma, double emin, double einterval,
                                   double species_mass, double gas_mass, double gas_density,
                                   const double *sigma, double *prob);
/* ProbilityReactive (MCCInit.F90:272-309); returns CollisionRatio */
double ora_probability_reactive(int nreaction, int nsigma, double species_mass, double gas_mass,
                                double gas_density, double beta_max, double polarizability,
                                double dt, const double *sigma, double *prob);

/* ProbilityHardShpere (MCCInit.F90:172-270) for nreaction 2 (101, 103) or 1
 * (101); returns SigmaMax */
double ora_probability_hard_sphere(int nreaction, int nsigma, double emin, double einterval, double species_mass,
                                   double gas_mass, double gas_density, double r_species, double r_gas,
                                   double *prob);

#ifdef __cplusplus
}
#endif
#endif
