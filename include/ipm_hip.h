/*
 * ipm_hip.h -- C ABI of the MI355X (gfx950) particle hot path for iPM2D.
 *
 * Every entry point returns 0 on success or a negative IPM_E* code; the text
 * of the last error is available from ipm_last_error().  Plain pointers and
 * sizes only, so Fortran (iso_c_binding), C, C++ and Python (ctypes) bind it
 * directly.  Grid arrays use the reference's Fortran column-major layouts
 * (z fastest) including the ghost layers the reference allocates:
 *   Ez(0:Nz, 1:Nr)   (Nz+1)*Nr doubles     R/src/pic/field/FieldEM.f90:72
 *   Er(1:Nz, 0:Nr)   Nz*(Nr+1) doubles     R/src/pic/field/FieldEM.f90:73
 *   Rho/RhoOne/ChiOne/sigma(1:Nz, 1:Nr)    Nz*Nr doubles
 *   Chi(0:Nz+1, 0:Nr+1)                    (Nz+2)*(Nr+2) doubles
 *                                          R/src/pic/field/FieldSource.f90:50-53
 * Particles are the 8 ParticleOne fields Z R Vz Vr Vt Az Ar WQ
 * (R/src/pic/particle/ParticleOne.f90:8) in grid units, resident on the GPU
 * as structure-of-arrays FP64.  Species are indexed 0..ns-1 as in the
 * reference (0 = electrons); gases 0..ng-1 (reference gas i+1).
 */
#ifndef IPM_HIP_H
#define IPM_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPM_OK 0
#define IPM_EINVAL (-1)    /* bad argument / index / shape */
#define IPM_EHIP (-2)      /* HIP runtime error */
#define IPM_ECAPACITY (-3) /* particle capacity exceeded (the reference stops:
                              R/src/pic/particle/ParticleBundle.f90:587-590) */
#define IPM_ENODEV (-4)    /* no usable GPU */

#define IPM_LAYOUT_AOS 0   /* n records of 8 doubles (the reference's PO(:) array) */
#define IPM_LAYOUT_SOA 1   /* 8 consecutive arrays of n doubles */

#define IPM_NS_MAX 9       /* the reference's ns_max (R/config.json:83) */
#define IPM_NG_MAX 3
#define IPM_NR_MAX 128     /* reactions per (species, gas) table */
/* Largest capacity of one species (2^29 - 32 particles, 4 GiB per field):
 * the particle pass addresses each field with 32-bit byte offsets. */
#define IPM_CAPACITY_MAX 536870880LL

typedef struct ipm_ctx ipm_ctx;

/* Context: one simulation on the current HIP device.  nz, nr = node counts
 * (Domain%GlobalShape), dx = dz = dr (R/src/base/init.F90:131-138), dt. */
int ipm_create(ipm_ctx **ctx, int nz, int nr, double dx, double dt, int ns, int ng);
int ipm_destroy(ipm_ctx *ctx);
const char *ipm_last_error(void);

/* Species parameters: ParticleBundle%Charge/%Mass (ParticleBundle.f90:350-351)
 * and the device capacity (npar_max, ParticleBundle.f90:354-359), at most
 * IPM_CAPACITY_MAX. */
int ipm_set_species(ipm_ctx *ctx, int s, double charge, double mass, int64_t capacity);

/* Geometry for absorption and deposit (R/src/base/Geometry.F90:21-45,
 * R/src/base/init.F90:402-424): inv_volume = point_inv_volume_weighting
 * (Nz*Nr); cell_material (Nz-1)*(Nr-1); hedge_type Nz*(Nr-1);
 * vedge_type (Nz-1)*Nr; material_type[m + ndiel] for m in [-ndiel, nmetal]
 * (R/src/pic/materials/Materials.F90:52-56). */
int ipm_set_geometry(ipm_ctx *ctx, const double *inv_volume, const int *cell_material,
                     const int *hedge_type, const int *vedge_type,
                     int ndiel, int nmetal, const int *material_type);

/* Particle transfer (restart Dump/Load, ParticleBundle.f90:117-150, and the
 * explicit host sync the reference's other per-particle readers need).
 * An upload clears the device's pending error reports (capacity, queue),
 * so a context recovers from IPM_ECAPACITY by re-uploading its species. */
int ipm_upload_species(ipm_ctx *ctx, int s, int64_t n, const double *po, int layout);
/* Copies n particles that already live in device memory -- 8 SoA arrays
 * with leading dimension ld, field f of particle i at po_dev[f*ld + i] --
 * into slots [offset, offset + n) of species s, and sets its count to
 * offset + n (slots below offset keep what an earlier call put there).
 * No counterpart in the reference: for callers whose particles are made or
 * kept on the GPU, such as device-side loading of runs too large to build
 * on the host.  Clears the pending error reports like ipm_upload_species. */
int ipm_upload_species_device(ipm_ctx *ctx, int s, int64_t offset, int64_t n, const double *po_dev, int64_t ld);
/* Downloads the live particles (dead slots dropped) into po (capacity cap). */
int ipm_download_species(ipm_ctx *ctx, int s, int64_t *n, double *po, int64_t cap, int layout);
/* Live particle count (ParticleBundle%NPar). */
int ipm_count(ipm_ctx *ctx, int s, int64_t *n);

/* FieldEM%Ez/%Er after SetE (R/src/pic/field/FieldSource.f90:95-120). */
int ipm_set_fields(ipm_ctx *ctx, const double *Ez, const double *Er);

/* PB(s)%MoveES(FG) -> MoveESIm (R/src/pic/particle/ParticleBundle.f90:649-748);
 * s = -1 runs every species. */
int ipm_move_es(ipm_ctx *ctx, int s);
/* MaterialAborptionParticle(CF, Geom, MT, PB(s))
 * (R/src/pic/particleboundary/ParticleMaterial.F90:10-90); s = -1: all. */
int ipm_absorb(ipm_ctx *ctx, int s);
/* Both of the above in one pass (OneStep's loop body, R/src/onestep.F90:141-144). */
int ipm_move_absorb(ipm_ctx *ctx, int s);
/* Adds the absorption tallies accumulated since the last call to
 * MT%metls(1:nmetal)%charge_one_step (metal_charge) and MT%sigma_one_step
 * (sigma, Nz*Nr), then clears them on the device.  Either may be NULL. */
int ipm_take_tallies(ipm_ctx *ctx, double *metal_charge, double *sigma);

/* FO(s)%P2CES(PB(s), Geom) (R/src/pic/field/FieldOne.f90:132-180):
 * rho_one/chi_one (Nz*Nr each) receive RhoOne/ChiOne; NULL keeps them on
 * the device only.  s = -1 deposits every species. */
int ipm_p2c(ipm_ctx *ctx, int s, double *rho_one, double *chi_one);
/* FT%Sum(FO) (R/src/pic/field/FieldSource.f90:72-92): rho (Nz*Nr) and
 * chi with ghosts ((Nz+2)*(Nr+2); ghosts set to 0 as FT%Zero does). */
int ipm_field_sum(ipm_ctx *ctx, double *rho, double *chi);

/* MCC cross-section bundle MCCBundle(s, g) (R/src/mc/TypeMCC.f90:23-33),
 * as built by MCCBundleInit (R/src/mc/MCCInit.F90:15-309).
 * rtype/resultant/threshold: per reaction; prob: Probility(j,i) column-major
 * (reaction fastest), nreaction*nsigma doubles. emax = 10**EnergyMax.
 * gas_mass [kg], gas_temperature [K], beta_max: the gas's GasOne fields. */
int ipm_set_mcc(ipm_ctx *ctx, int s, int g, int model, int nreaction, int nsigma,
                double emin, double einterval, double emax, double collision_ratio,
                const int *rtype, const int *resultant, const double *threshold,
                const double *prob, double gas_mass, double gas_temperature, double beta_max);
/* Counter-based random stream (Philox4x32-10) replacing the reference's
 * unseeded random_number: draws are keyed by (seed, step, species|gas<<8,
 * slot, draw).  step advances by one after every MCC pass.  Also clears the
 * device's pending error reports, like ipm_upload_species. */
int ipm_set_rng(ipm_ctx *ctx, uint64_t seed, uint32_t step);
/* MCC(Ns, Ng, PB, SO, GO, MCCB) (R/src/mc/MCCPublic.F90:38-121) including
 * MCCAddNewParticle (209-226): products are appended to their species. */
int ipm_mcc(ipm_ctx *ctx);
/* Enables (on != 0) or disables the per-event counters below (off by default:
 * counting adds one atomic per collision). */
int ipm_set_event_counting(ipm_ctx *ctx, int on);
/* Null-event screen (on by default; no counterpart in the reference, whose
 * MCCPublic.F90:61-97 runs Updater and SelectProbility for every selected
 * particle): a selected particle whose reaction-choice uniform is at or above
 * every probability its table can reach over the relative energies the
 * target sampler allows is a null event, known without sampling the target,
 * and is never queued for the collision pass.  Exact: results are identical
 * with the screen on or off (off: every selection takes the full path). */
int ipm_set_null_screen(ipm_ctx *ctx, int on);
/* Per-(species, gas, reaction index) event counts since the last call
 * (counts[(s*ng + g)*128 + j], j = 0 for a null event); cleared after read. */
int ipm_take_mcc_counts(ipm_ctx *ctx, int64_t *counts);

/* The whole per-step particle path in one persistent launch over every
 * species: move + absorb + deposit + MCC (+ grid epilogue producing Rho and
 * Chi).  Equivalent to the reference's OneStep particle calls
 * (R/src/onestep.F90:141-177) because MCC reads no field: the Poisson solve
 * that the reference runs between deposit and MCC does not feed it. */
int ipm_step(ipm_ctx *ctx);
/* Same step replayed from a captured hipGraph (captured on first use). */
int ipm_step_graph(ipm_ctx *ctx);
/* ipm_step and ipm_step_graph only enqueue work: errors they raise on the
 * device (capacity overflow) are returned by the next synchronizing call
 * (ipm_synchronize, ipm_count, downloads, ipm_field_sum, ...). */

/* HIP-event timing of the kernels ipm_step enqueues (eager mode only):
 * ms[0..2] / launches[0..2] = summed milliseconds and launch counts of the
 * particle pass, the collision pass and the grid epilogue since the last
 * call; cleared after read. */
int ipm_set_timing(ipm_ctx *ctx, int on);
int ipm_take_timing(ipm_ctx *ctx, double *ms, int64_t *launches);
/* Split/merge of species s inside every later ipm_step / ipm_step_graph:
 * split_particle_weight and merge_particle_random (R/src/base/Rezoning.F90:
 * 53-92) with the arguments of ipm_split_merge, applied in the fused
 * particle pass where OneStep calls particle_split_merge_control
 * (R/src/onestep.F90:160): after the deposit, before the collision pass,
 * which also draws for the copies.  mode 0 switches it off again.  Weight
 * and the other parameters may change between steps (the reference updates
 * Weight every CF%Period) without a new graph.  Results equal ipm_move_absorb,
 * ipm_p2c, ipm_split_merge and ipm_mcc called in that order. */
int ipm_set_step_rezone(ipm_ctx *ctx, int s, int mode, double weight, double alpha, double grid_r,
                        double merge_limit);
/* Cell sort (counting sort by cell; drops dead slots). s = -1: all.  No
 * counterpart in the reference (its bundles are unordered); nothing the
 * path computes depends on particle order. */
int ipm_sort(ipm_ctx *ctx, int s);
/* Sorts every species after each `steps` calls of ipm_step / ipm_step_graph
 * (0, the default: never).  ipm_step_graph then replays the sort as a
 * captured graph of its own, so a run of many steps needs no host
 * synchronization at all.  A scheduling choice: results do not depend on it
 * (the sort changes storage order only). */
int ipm_set_sort_interval(ipm_ctx *ctx, int steps);
/* Slots per tile of the cell sort: 2048 (256 threads, the default) or 4096
 * (512 threads).  A tile's particles of one cell go to consecutive slots, so
 * a larger tile writes longer runs.  A scheduling choice with no counterpart
 * in the reference (which does not sort): the sorted set, its cell order and
 * every later result do not depend on it. */
int ipm_set_sort_tile(ipm_ctx *ctx, int slots);

/* Launch shape of the fused particle pass (ipm_step, ipm_step_graph):
 * block = threads per block of a tile of 2048 slots, 128, 256 or 512 (16, 8
 * or 4 particles per thread; 0 = 256); blocks_per_cu = 0 launches one block
 * per tile (grid sized from the capacities, the default), 1..64 a
 * persistent grid of that many blocks per compute unit looping over the
 * tiles.  A scheduling choice with no counterpart in the reference: results
 * do not depend on it. */
int ipm_set_launch(ipm_ctx *ctx, int block, int blocks_per_cu);
/* Height (nodes in r: 16, 24 or 32, default 24) of the LDS window the fused
 * pass deposits into before flushing to HBM (64 nodes in z); heights other
 * than 24 run with 256-thread blocks.  Results do not depend on it. */
int ipm_set_deposit_window(ipm_ctx *ctx, int rows);

/* Grid of the collision pass: blocks_per_cu (1..6, default 6) blocks per
 * compute unit.  Like ipm_set_launch, a scheduling choice with no
 * counterpart in the reference: results do not depend on it. */
int ipm_set_collide_grid(ipm_ctx *ctx, int blocks_per_cu);
int ipm_synchronize(ipm_ctx *ctx);

/* Device addresses of species s's 8 SoA arrays (for zero-copy interop) and
 * of the raw per-species deposit grid. */
int ipm_species_device_ptrs(ipm_ctx *ctx, int s, double **fields8);

/* FT%SetE(FG) on the device (R/src/pic/field/FieldSource.f90:95-120):
 * phi(1:Nz, 1:Nr) is the node potential; its ghost layers are filled by
 * linear extrapolation as the reference's field communicator does for every
 * side of a one-rank domain (R/src/pic/domain/PICCom.F90:474-590), and Ez,
 * Er are formed on the device.  Replaces ipm_set_fields when the host has
 * the potential: one grid crosses PCIe instead of two. */
int ipm_set_potential(ipm_ctx *ctx, const double *phi);

/* Geometry of the field solve (R/src/base/Geometry.F90:21-45):
 * cell_epsilon = Geom%cell_epsilon ((Nz-1)*(Nr-1) relative permittivities)
 * and point_type = Geom%point_type (Nz*Nr, > 0 on metal nodes).  Needs
 * Nz, Nr >= 3. */
int ipm_set_solver_geometry(ipm_ctx *ctx, const double *cell_epsilon, const int *point_type);
/* The linear system of the host Poisson solve, assembled on the device from
 * the Rho and Chi of the last ipm_field_sum / ipm_step / ipm_step_graph:
 * FS%SetCoe(FT, Geom) (with_chi != 0) or FS%SetCoeWN(Geom) (with_chi == 0),
 * then FS%SetSrc(FT, MT, Geom) and FS%SetBound(boundary)
 * (R/src/pic/field/FieldSolver.f90:136-301).
 *   coe    (5*Nz*Nr, NULL to skip): CoeA, CoeB, CoeC, CoeD, CoeE, each Nz*Nr,
 *          the five-point coefficients of phi(i-1,j), phi(i,j), phi(i+1,j),
 *          phi(i,j-1), phi(i,j+1) at node (i, j);
 *   source (Nz*Nr, NULL to skip): -dx^2/Epsilon * (Rho + sigma *
 *          point_inv_volume_weighting), then bound where bound_mask != 0;
 *   sigma  (Nz*Nr) = MT%sigma, NULL for zero; bound (Nz*Nr doubles) and
 *          bound_mask (Nz*Nr ints) = boundary%bound / %bound_mask, both
 *          NULL to merge nothing.
 * These are the arrays poisson_solve_ receives (FieldSolver.f90:307-316):
 * the host's pass over the grid is replaced by one device kernel. */
int ipm_solver_system(ipm_ctx *ctx, int with_chi, const double *sigma, const double *bound, const int *bound_mask,
                      double *coe, double *source);

/* Moment diagnostics of species s: WeightingParticleField
 * (R/src/diag/diagParticleField.F90:173-318).  Each output is Nz*Nr (or
 * NULL to skip): rho = RhoOne (number density), energy_ev = TOne (mean
 * energy in eV where rho > 0), mp = MPOne (macro-particles per node),
 * jz/jr = JzOne/JrOne (A/m^2).  implicit_j != 0 takes the current from a
 * copy pushed once more in the present field, as the reference's implicit
 * scheme does (diagParticleField.F90:225-236; needs ipm_set_fields);
 * implicit_j == 0 uses the particle's own velocity. */
int ipm_moments(ipm_ctx *ctx, int s, int implicit_j, double *rho, double *energy_ev, double *mp, double *jz,
                double *jr);
/* ipm_moments for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): no counterpart in the reference.
 * Allocates nothing -- the sums live in a static device array of at most
 * 2^20 grid nodes (40 MiB per GPU, held until the process ends); larger
 * grids return IPM_EINVAL.  The outputs are DEVICE arrays of Nz*Nr doubles
 * (NULL to skip); the work is only enqueued on the library's stream, so
 * read them after ipm_synchronize.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU race. */
int ipm_moments_noalloc(ipm_ctx *ctx, int s, int implicit_j, double *rho, double *energy_ev, double *mp,
                        double *jz, double *jr);

/* Weight rezoning of species s: split_particle_weight then
 * merge_particle_random (R/src/base/Rezoning.F90:53-92), the per-bundle
 * calls of particle_split_merge_control (Rezoning.F90:17-50).  The reference
 * weight of a particle at R is CalRW (R/src/pic/particle/ParticleBundle.f90:393-402)
 *   ref = weight * ((1 - alpha) * R / grid_r + alpha),
 * with weight = PB%Weight, alpha = PB%alpha, grid_r = PB%grid_r = Nr - 1.
 * mode & 1 splits: Ncr = int(WQ / ref); Ncr > 1 appends Ncr - 1 copies of
 * weight ref and subtracts ref from WQ once per copy.  mode & 2 merges: WQ <
 * merge_limit * ref keeps the particle at weight ref when a uniform draw is
 * below WQ / ref and deletes it otherwise.  Draws come from the particle's
 * state-keyed stream under a tag of their own, so the result does not depend
 * on storage order.  Copy j of a split differs from its parent only in the
 * low bits of Vt's bit pattern (XOR j), so that copies draw independently
 * afterwards.  stats (3 entries, may be NULL) receives copies appended,
 * merge candidates kept and merge candidates deleted.  Returns
 * IPM_ECAPACITY (nothing split for the particles concerned) when copies
 * would exceed the species capacity; the reference grows its array. */
int ipm_split_merge(ipm_ctx *ctx, int s, int mode, double weight, double alpha, double grid_r,
                    double merge_limit, int64_t *stats);

/* Volume recombination: RecombinationLocal%recb(Ns, PB, time_step), i.e.
 * executingRecombination2D + deletingParticles (R/src/mc/Recombination.F90:164-273),
 * with the reactions initRecombination2D stores (Recombination.F90:43-140).
 * reaction_list is recb_reaction(nreaction, 3) in Fortran order with 0-based
 * species: first species of reaction i at [i], second at [nreaction + i],
 * type at [2*nreaction + i] (0 constant rate, 1 temperature form); rate[i]
 * is recb_rate(i); time_step = recombination_time_step_number * dt.  Cell
 * densities (and inverse mean energies for type 1 species) are formed once,
 * then each reaction removes its budget from both species' particles of
 * every cell, visiting them from the highest storage slot down; the draw is
 * the particle's state-keyed uniform.  Each involved bundle is first
 * reordered by cell, keeping the slot order inside a cell and dropping dead
 * slots; deleted particles become dead slots.  deleted (ns entries, may be
 * NULL) receives the macro-particles deleted per species. */
int ipm_recombine(ipm_ctx *ctx, int nreaction, const int *reaction_list, const double *rate, double time_step,
                  int64_t *deleted);

/* ---- Differentiable collisionless window ----------------------------------
 * No counterpart in the reference (which has no gradients).  Advances a
 * snapshot of the resident particles through nsteps steps of MoveESIm
 * (R/src/pic/particle/ParticleBundle.f90:671-748) and the absorption rule of
 * MaterialAborptionParticle (R/src/pic/particleboundary/ParticleMaterial.F90:22-41)
 * in the field E = E0 + coef[n] * E1 of step n, with a loss per step
 *   loss[n] = sum_i alpha[s_i] * WQ_i * Vz_i * M(Z_i, R_i),
 *   M = P2CES-weighted interpolation (R/src/pic/field/FieldOne.f90:151-166)
 *       of the node field mu (Nz*Nr),
 * taken after the push and absorption of step n.  With mu = mask *
 * point_inv_volume and alpha[s] = dt * w_s * Charge_s * VFactor, loss[n] =
 * dt * sum_s w_s * sum_nodes mask * JzOne_s (explicit current, see
 * ipm_moments).  No collisions happen inside the window.
 *
 * ipm_window_create snapshots every resident slot of every species (sort
 * first: the adjoint skips whole wavefronts whose particles never reach the
 * support of mu, which cell order makes common) and keeps a checkpoint of
 * the state every `interval` steps (1 <= interval <= 64).  E0/E1 use the
 * Ez/Er layouts above; alpha has ns entries. */
int ipm_window_create(ipm_ctx *ctx, int nsteps, int interval, const double *Ez0, const double *Er0,
                      const double *Ez1, const double *Er1, const double *mu, const double *alpha);
/* Forward over the window from the snapshot; loss receives nsteps values.
 * record != 0: particles are deleted by the reference's absorption rule and
 * each one's deletion step is stored; record == 0 replays the stored
 * deletion steps (events frozen), so the loss is a smooth function of coef. */
int ipm_window_forward(ipm_ctx *ctx, const double *coef, int record, double *loss);
/* Adjoint: dcoef[n] = d(sum_m gloss[m] * loss[m]) / d coef[n] for the
 * replayed window at coef (checkpoints are rebuilt first when the last
 * forward used other coefficients). */
int ipm_window_backward(ipm_ctx *ctx, const double *coef, const double *gloss, double *dcoef);
/* Particles in the window, checkpoint segments, and HIP-event milliseconds
 * of the last forward and the last adjoint. */
int ipm_window_info(ipm_ctx *ctx, int64_t *nparticles, int *nsegments, float *forward_ms, float *adjoint_ms);
/* Checkpoint k (0 = snapshot, nsegments = end of window) as 7 arrays of
 * nparticles doubles (Z R Vz Vr Vt Az Ar), and each particle's deletion step
 * (nsteps = never, -1 = empty slot).  Either pointer may be NULL. */
int ipm_window_download(ipm_ctx *ctx, int checkpoint, double *state, int *absorbed_step);
int ipm_window_destroy(ipm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
