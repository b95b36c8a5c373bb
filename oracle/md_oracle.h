/*
 * md_oracle.h -- FP64 CPU restatement of the MD_Brushes force-and-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product never links or calls it.
 *
 * Every function restates one reference routine (file:line cited at the
 * definition in md_oracle.c), in plain serial C, compiled with
 * -ffp-contract=off so that sums of products round as the reference writes
 * them.  Arrays follow the reference's Fortran column-major layout:
 * pos(NN,4), vel(NN,3), acc(NN,3), eta(NN,3).
 */
#ifndef MD_ORACLE_H
#define MD_ORACLE_H

#include <stdint.h>
#include "../include/md_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

int orc_derive(md_params *p);
int orc_build_topology(const md_params *p, int32_t *charge, int32_t *fene_list,
                       int32_t *anchor_list);

orc_sys *orc_create(const md_params *p);
void orc_destroy(orc_sys *s);
void orc_set_state(orc_sys *s, const double *pos, const double *vel,
                   const double *acc);
void orc_get_state(orc_sys *s, double *pos, double *vel, double *acc,
                   double *acc_c, double *dr_max);
void orc_set_drmax(orc_sys *s, double d1, double d2);

int64_t orc_lj_verlet_list(orc_sys *s);
int64_t orc_real_verlet_list(orc_sys *s);
int64_t orc_get_list(orc_sys *s, int which, int32_t *pairs, int64_t cap);

int orc_lj_force(orc_sys *s, double *f);          /* LJ + wall, NN*3       */
int orc_fene_force(orc_sys *s, double *f);        /* FENE, NN*3; 2 = break */
void orc_real_space(orc_sys *s, double *f);       /* real space + slab     */
void orc_spme(orc_sys *s, double *f);             /* SPME reciprocal       */
void orc_standard_ewald(orc_sys *s, double alpha, int k1, int k2, int k3,
                        double *f);
void orc_error_analysis(orc_sys *s, double out[8], double *acc_se, double *acc_c);
int orc_compute_force(orc_sys *s, int32_t step, int32_t noise,
                      const double *eta);
int orc_new_position(orc_sys *s, int32_t step, int32_t noise,
                     const double *eta);
int orc_update_verlet_list(orc_sys *s);
void orc_rescale_velocity(orc_sys *s);
int orc_run(orc_sys *s, int32_t step0, int32_t nsteps, int32_t preheat);
void orc_energies(orc_sys *s, double *out);  /* out[16], see md_oracle.c */
void orc_set_threads(int n);                 /* OpenMP build: threads of the hot loops */

/* building blocks exposed for tests */
void orc_fft3d(int n1, int n2, int n3, double *re_im, int sign);
void orc_bspln_coeffs(int ord, double *cof, double *dcof);
double orc_bspln(double x, int ord);
void orc_splcof(int ord, const double *w_in, int n, const double *cof,
                const double *dcof, double *sp, double *dsp);
void orc_pme_bc_axis(int ng, int ord, double gdim, double alpha, double *b2,
                     double *mv, double *ev);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
void orc_noise_normals(uint64_t seed, int32_t step, int32_t i, double n[3]);

#ifdef __cplusplus
}
#endif
#endif
