/*
 * md_engine.h -- C-ABI of the MI355X force-and-step engine for Langevin
 * dynamics of grafted polyelectrolyte brushes (linear, star and mixed).
 *
 * The reference program keeps all state in Fortran module variables and its
 * hot-path routines take no arguments.  This header is the argument-passing
 * form of those routines; each entry point names the reference routine it
 * replaces (file:line in wangshaoyun/MD_Brushes).  A Fortran host binds it
 * with iso_c_binding; Python binds it with ctypes.
 *
 * Conventions
 *   - Arrays are the reference's Fortran arrays in column-major order:
 *     pos(NN,4) = x(1:NN), y(1:NN), z(1:NN), q(1:NN); vel(NN,3); acc(NN,3).
 *   - Particle indices in topology arrays are 1-based, as in the reference.
 *   - Every int-returning call returns MD_OK (0) or an MD_E* status; the
 *     message is available from md_last_error().
 *   - All device state is owned by the engine and lives until md_destroy().
 *     One host thread per engine; calls on one engine are not re-entrant.
 */
#ifndef MD_ENGINE_H
#define MD_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MD_MAX_ARM 16

enum {
  MD_OK = 0,
  MD_EINVAL = 1,        /* bad argument or parameter                       */
  MD_EBOND = 2,         /* FENE bond stretched past R0 (compt_acc.f90:879)   */
  MD_ECELL = 3,         /* particle outside the cell grid (x==Lx/2, z<0, z>=Lz) */
  MD_ENONFINITE = 4,    /* non-finite force                                 */
  MD_EOVERFLOW = 5,     /* neighbour-list or cell capacity exceeded          */
  MD_EHIP = 6,          /* HIP runtime error                                 */
  MD_ENOMEM = 7,        /* device allocation failed                          */
  MD_EIO = 8            /* parameter file could not be read                  */
};

enum { MD_FP64 = 0, MD_FP32 = 1 };

/* Noise source for compute_force / new_position. */
enum {
  MD_NOISE_NONE = 0,    /* deterministic forces only                        */
  MD_NOISE_ETA = 1,     /* injected standard normals eta(NN,3)              */
  MD_NOISE_PHILOX = 2   /* counter-based Philox4x32-10 stream keyed by seed  */
};

/*
 * Parameters.  The first two blocks are the values of system_data.txt
 * (in_output.f90:189-213) and force_data.txt (compt_acc.f90:140-157) in file
 * order; charged_arm is q_arm.txt (in_output.f90:241-247).  md_derive_params
 * fills the derived block exactly as the reference derives it
 * (in_output.f90:253-288, compt_acc.f90:159-162, compt_acc.f90:375-420).
 * The override block lets a caller fix the SPME mesh and Ewald alpha instead
 * of the reference's automatic rule; zero means "reference rule".
 */
typedef struct md_params {
  /* system_data.txt */
  int32_t uniform_or_random;
  int32_t arm, Nma, Nga, Nml, Ngl, man_l, man_s;
  int32_t StepNum0, StepNum, DeltaStep1, DeltaStep2, DeltaStep3, multistep;
  double Lz, Z_empty, sigmag, Beta, qq, qqi, ion_ratio, R_bond, dt;
  /* force_data.txt */
  double epsilon, sigma, sigma_s, rcl, rvl, rsk, R0_2, kFENE;
  int32_t ordr[3];
  int32_t pad0;
  double lb, xi, EF, tol, tau_rf;
  /* q_arm.txt */
  int32_t charged_arm[MD_MAX_ARM];
  /* overrides (0 = reference rule) */
  int32_t ovr_ng[3];
  int32_t pad1;
  double ovr_alpha;
  double ovr_rcc;
  /* derived */
  int32_t Nta, Ntl, Npe, N_anchor, Nq_PE, Nq_salt_ions, Nq, NN, N_bond;
  int32_t real_verlet;
  int32_t ng[3];
  int32_t pad2;
  double Lx, Ly, Zfac;                 /* Zfac: Z_empty after in_output.f90:288 */
  double sigma_sp, rcl_s, rcl_sp, rvl_eff;
  double alpha, rcc, rvc;
  double gdim[3];
  /* engine options */
  int32_t precision;                   /* MD_FP64 or MD_FP32                     */
  int32_t pad3;
  uint64_t seed;                       /* key of the Philox noise stream         */
  /* engine geometry (0 = the reference's; md_autotune_geometry chooses) */
  double ovr_skin;        /* LJ list: rvl_eff = rcl_s + 2 ovr_skin (reference rcl_s + 1)  */
  double ovr_cell_scale;  /* real-space sweep cells of edge >= ovr_cell_scale * r_c (>= 1) */
  int32_t ovr_tile[3];    /* spreading tile edge per axis, 4..16 mesh points (16)         */
  int32_t pad4;
} md_params;

typedef struct md_engine md_engine;

/* Replaces read_data (in_output.f90:171-214), read_force_parameters
 * (compt_acc.f90:138-157) and the q_arm.txt read (in_output.f90:241-247).
 * q_arm may be NULL when Nga == 0.  Derived fields are left for
 * md_derive_params. */
int md_read_params(const char *system_data, const char *force_data,
                   const char *q_arm, md_params *p);

/* Replaces data_operation (in_output.f90:217-288), the derived radii of
 * read_force_parameters (compt_acc.f90:159-162) and
 * Initialize_ewald_parameters (compt_acc.f90:375-430). */
int md_derive_params(md_params *p);

/* Replaces build_fene_list (compt_acc.f90:317-372) and build_charge
 * (compt_acc.f90:433-488).  Outputs are 1-based:
 * charge(Nq), fene_list(N_bond,2) column-major, anchor_list(N_anchor). */
int md_build_topology(const md_params *p, int32_t *charge, int32_t *fene_list,
                      int32_t *anchor_list);

/* Seeded restatement of Initialize_position (init_update.f90:6-66) with
 * uniform or random grafting (init_update.f90:69-394) and ions
 * (init_update.f90:397-458); pos(NN,4) column-major.  The reference seeds
 * from the OS (main.f90:12); here the stream is fixed by `seed`. */
int md_initialize_position(const md_params *p, uint64_t seed, double *pos);

/* Replaces initialize_velocity (init_update.f90:462-495): Gaussian
 * velocities with variance 1/Beta, anchors at rest, then rescale_velocity. */
int md_initialize_velocity(const md_params *p, uint64_t seed, double *vel);

/* Allocates device state for one system (initialize_force_parameters,
 * compt_acc.f90:85-135: B-spline coefficients, PME_BC table, buffers).
 * The topology is built from p with md_build_topology. */
int md_create(md_engine **e, const md_params *p);

/* Uploads pos(NN,4), vel(NN,3), acc(NN,3) (acc may be NULL = zero). */
int md_upload_state(md_engine *e, const double *pos, const double *vel,
                    const double *acc);

/* Downloads state; any pointer may be NULL.  dr_max[2] = (dr_max1, dr_max2). */
int md_download_state(md_engine *e, double *pos, double *vel, double *acc,
                      double *dr_max);

/* Held Coulomb force acc_c(NN,3) (compt_acc.f90:55). */
int md_download_coulomb(md_engine *e, double *acc_c);

/* Replaces lj_verlet_list (compt_acc.f90:185-314) and, when real_verlet
 * is set, real_verlet_list (compt_acc.f90:491-626).  The real-space list's
 * capacities are sized at the first call (a host scan of the positions);
 * the step loop never needs that list, so uploads skip the scan. */
int md_build_lists(md_engine *e);

/* Verlet lists as (i,j) pairs, 1-based, sorted by (i,j).  Returns the pair
 * count in *npair; writes at most cap pairs into pairs(cap,2) column-major.
 * which = 0: LJ list, 1: real-space list. */
int md_download_list(md_engine *e, int32_t which, int32_t *pairs, int64_t cap,
                     int64_t *npair);

/* Replaces Compute_Force (compt_acc.f90:815-862) for step `step`.
 * noise: MD_NOISE_*; eta(NN,3) standard normals when noise == MD_NOISE_ETA. */
int md_compute_force(md_engine *e, int32_t step, int32_t noise,
                     const double *eta);

/* Replaces new_position (init_update.f90:555-607) including
 * period_condition_pos (init_update.f90:610-643). */
int md_new_position(md_engine *e, int32_t step, int32_t noise,
                    const double *eta);

/* Replaces update_verlet_list (compt_acc.f90:699-726). */
int md_update_verlet_list(md_engine *e);

/* Replaces rescale_velocity (init_update.f90:498-552). */
int md_rescale_velocity(md_engine *e);

/* Runs steps step0 .. step0+nsteps-1 of the loop at main.f90:67-94
 * (new_position, then in pre-heat the rescale every 20 steps below 300000,
 * then update_verlet_list) with Philox noise, on the device, without host
 * round trips.  preheat selects the main.f90:67-75 form. */
int md_run(md_engine *e, int32_t step0, int32_t nsteps, int32_t preheat);

/* Deterministic force parts for accuracy checks: writes
 * f_short(NN,3) = LJ + wall + FENE, f_coul(NN,3) = SPME + real space + slab,
 * evaluated at the current positions (no noise, no field, no anchor mask). */
int md_force_parts(md_engine *e, double *f_short, double *f_coul);

/* error_analysis (compt_acc.f90:729-812) on the device: the accuracy of the
 * engine's Coulomb force against a standard-Ewald reference computed in FP64
 * from the resident positions.  The reference force follows the routine:
 * Initialize_ewald_parameters at tol = 5, tau_rf = 10 (no overrides),
 * k-ranges ceiling(Kmax/3.), Standard_Ewald + real_space (every pair within
 * r_c, x,y minimum image, slab correction).  out[8]: rmsf =
 * sqrt(sum |F_se - F_c|^2 / sum |F_se|^2) over every particle (the reference
 * sums rows 1..Nq), the Kolafa-Perram estimates f_r and f_k of the engine's
 * own parameters, and the reference's alpha, r_c, K1, K2, K3.  acc_se (NN,3),
 * may be null: the reference force.  Positions, velocities, forces and the
 * held Coulomb force are left as they are. */
int md_error_analysis(md_engine *e, double *out, double *acc_se);

/* Mean duration in milliseconds of each kernel of a step, timed with HIP
 * events on the engine's stream over `iters` launches (no reference
 * counterpart: measurement only; short_range is the launch md_run's
 * replayed steps issue (list bookkeeping, second half kick, then the next
 * step's half kick and drift into the other position buffer), on a step
 * without a list rebuild, over 8 back-to-back launches per iteration, each
 * reading the buffer the one before wrote; kick_drift, the form that starts
 * a run, as the difference between it followed by those launches and the
 * launches alone).
 * ms[15]: kick_drift, short_range, bin_q, sort_q, spread, fft_x_fwd,
 * fft_y_fwd, fft_z_conv, fft_y_inv, fft_x_inv, gather, real, bin_lj,
 * sort_lj, build_lj (with fused FFT planes, fft_x_* time the plane kernels
 * and fft_y_* are empty; sort_lj is empty, its rank pass is part of
 * build_lj).  Positions, velocities, forces, the list bookkeeping and the
 * rebuild queues are restored after each iteration; the lists are rebuilt
 * and the held Coulomb force is recomputed at the current positions. */
int md_time_kernels(md_engine *e, int32_t iters, double *ms, int32_t n);

/* Chooses the engine's launch forms by timing its Coulomb evaluation under
 * each (no reference counterpart: tuning only): fused FFT planes or
 * separate x/y passes, z-pass lines per workgroup, a wavefront or a
 * workgroup per real-space work item, the real-space sweep in a launch of
 * its own or inside the spreading and gather launches, the sweep's
 * workgroup count, the spreading as its own launch or inside the forward
 * plane FFT kernel (k_plane_spread), how many of the sweep's workgroups
 * riding in the plane and gather launches are dispatched ahead of the
 * launch's own.  The forms give the same results.  choice[9] (may be null)
 * receives fused (0/1), z lines, workgroup (0/1), shared launches (0/1),
 * the workgroup form's grid cap (1024 or 2048 workgroups), spreading inside
 * the forward plane kernel (0/1), and with it the launches the sweep rides
 * in (1: gather, 2: plane and gather, 3: plane, z pass and gather), and the
 * riders dispatched first in the plane launch (0, 256, 512 or 768) and in
 * the gather launch (0, 512, 1024 or 65536: all).
 * Leaves the state as it was; the held Coulomb force is recomputed for the
 * current positions. */
int md_autotune(md_engine *e, int32_t iters, int32_t *choice);

enum { MD_TUNE_LISTS = 1 };

/* Engine geometry for a system (no reference counterpart: tuning only),
 * chosen by running it from pos(NN,4), vel(NN,3) in fresh engines: the
 * spreading tile edges ovr_tile (results unchanged) and, with flags &
 * MD_TUNE_LISTS, the LJ Verlet skin ovr_skin and the real-space sweep's
 * cell scale ovr_cell_scale (the same pairs within every cut in another
 * order: forces change by rounding; the reference's skin is 0.5).  Each
 * candidate runs `cycles` multistep cycles after ten warm-up cycles; knobs
 * are tuned one after the other and a candidate must beat the incumbent by
 * 1% (3% for the list knobs).  *best receives p with the chosen overrides,
 * derived; *steps_per_s (may be null) its measured rate. */
int md_autotune_geometry(const md_params *p, const double *pos, const double *vel,
                         int32_t cycles, int32_t flags, md_params *best,
                         double *steps_per_s);

/* ---- Replica batches (no reference counterpart: the reference runs one
 * system per process).  One engine holds many independent replicas of
 * uneven size and one precision; every kernel launch covers all of them, and
 * a replica's result does not depend on which others share the launch. */

/* Empty engine for replicas of the given precision and Coulomb interval. */
int md_batch_create(md_engine **e, int32_t precision, int32_t multistep);

/* Adds a replica with state pos(NN,4), vel(NN,3): builds its lists and its
 * step-0 force (Compute_Force at main.f90:37); its first step is 1.
 * *slot is a stable handle. */
int md_batch_add(md_engine *e, const md_params *p, const double *pos,
                 const double *vel, int32_t *slot);

/* Frees a replica's device memory. */
int md_batch_remove(md_engine *e, int32_t slot);

/* Queues ncycles * multistep steps (new_position + update_verlet_list,
 * Philox noise) of every replica, the first cycle as a captured graph's
 * replay and the others as launch sequences on the engine's stream
 * (MD_GRAPH=1: every cycle replayed, MD_GRAPH=0: none), and returns
 * without waiting: md_batch_add, _status, _download and _observe
 * wait for them, so host work between calls overlaps the device. */
int md_batch_run(md_engine *e, int32_t ncycles);

/* Status of one replica (MD_OK or its device-side error) and its next step. */
int md_batch_status(md_engine *e, int32_t slot, int32_t *step, char *buf,
                    int32_t len);

/* State of one replica; any pointer may be NULL. */
int md_batch_download(md_engine *e, int32_t slot, double *pos, double *vel,
                      double *acc);

/* Device bytes held by one replica (slot >= 0) or by all (slot < 0), and the
 * replica count. */
int md_batch_info(md_engine *e, int32_t slot, int64_t *bytes, int32_t *nrep);

/* Device bytes that md_batch_add would allocate for a replica with these
 * parameters and positions pos(NN,4) (the same sizing, lists sized from the
 * configuration; the read-only tables shared by replicas of one mesh not
 * counted), computed without touching the device: what a request loop
 * admits against its HBM budget.  No reference counterpart. */
int md_replica_bytes(const md_params *p, const double *pos, int64_t *bytes);

/* Observables computed on the device: brush height (mean z of the polymer
 * beads, in_output.f90:1017-1020) and the z-density histogram of the polymer
 * beads, nbins = 500 bins of Lz/500 (in_output.f90:1878-1886). */
int md_batch_observe(md_engine *e, int32_t slot, double *height,
                     double *profile, int32_t nbins);

/* md_error_analysis for one replica of a batch (an accuracy spot check of a
 * served replica, compt_acc.f90:729-812); the replica's trajectory is not
 * changed. */
int md_batch_error_analysis(md_engine *e, int32_t slot, double *out,
                            double *acc_se);

/* md_batch_error_analysis that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its buffers live in a static device scratch array of 64 MiB
 * per GPU, held from library load to process end; a replica whose check
 * needs more is refused with MD_EINVAL.  The k-vectors are the whole box
 * with zero weight outside the reference's ellipsoid, so the results equal
 * md_batch_error_analysis's up to the grouping of the sums.  Calls that
 * overlap on one GPU share its scratch and race.  The engine must have run
 * a cycle since its last membership change (MD_EINVAL otherwise). */
int md_batch_error_analysis_noalloc(md_engine *e, int32_t slot, double *out,
                                    double *acc_se);

/* The height analysis of write_height (in_output.f90:1017-1079, columns 2-9
 * of height.txt, :1342-1344) from the resident positions: out[8] = h_avg
 * (polymer beads), hl_max, hl_end, hl_avg (linear chains: mean over chains
 * of the highest bead and of the free end; mean over their beads), hs_max,
 * hs_end, hs_branch, hs_avg (stars: mean of the highest bead, of the ends of
 * arms 2..arm, of the branch point; mean over their beads).  A molecule kind
 * the system lacks reads 0 (the reference divides by zero). */
int md_heights(md_engine *e, double *out);
int md_batch_heights(md_engine *e, int32_t slot, double *out);

/* Energies and pair virials at the current state, FP64 sums (pair and bond
 * arithmetic in the engine's precision).  The reference computes only the
 * kinetic energy (height, in_output.f90:1307-1311); the potentials are those
 * whose negative gradients are its forces: WCA 4((s/r)^12 - (s/r)^6) + 1 per
 * pair within the class cut (compt_acc.f90:922-952), wall 0.4/3 ((s/d)^9 -
 * (s/d)^3) (:955-975), FENE -k R0^2/2 ln(1 - r^2/R0^2) (:865-892), real
 * space lB/Beta q_i q_j erfc(a r)/r (:997-1038), reciprocal lB/Beta sum Q U /
 * (2 K1 K2 K3) (:1368-1398), slab 2 pi lB/Beta (sum q z)^2 / V (:1040-1048).
 * out[16]: WCA, wall, FENE, real space, reciprocal, slab, Ewald self term,
 * Coulomb (the four before), field -E_z sum q z, potential total, kinetic
 * sum |v|^2/2, kinetic / (NN - N_anchor) (the reference's kenetic_energy),
 * pair virials sum r.F of WCA, FENE and real space, 0.  The held Coulomb
 * force and the trajectory are unchanged. */
int md_energies(md_engine *e, double *out);
int md_batch_energies(md_engine *e, int32_t slot, double *out);

/* Exact resume (no reference counterpart: the reference restarts from its
 * six-decimal text checkpoint and rescales the velocities, main.f90:40-64).
 * md_state_save writes the engine's complete state between steps
 * (positions, velocities, forces, held Coulomb force, LJ list and cells,
 * accumulated displacements) into buf, md_state_size bytes; md_state_load
 * restores it into an engine made from the same parameters and precision,
 * after which the run continues bit for bit as if never interrupted. */
int md_state_size(md_engine *e, int64_t *bytes);
int md_state_save(md_engine *e, void *buf, int64_t cap);
int md_state_load(md_engine *e, const void *buf, int64_t len);

/* Copies the last error message (NUL-terminated) into buf. */
int md_last_error(md_engine *e, char *buf, int32_t len);

void md_destroy(md_engine *e);

#ifdef __cplusplus
}
#endif

#endif /* MD_ENGINE_H */
