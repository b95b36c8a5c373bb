/*
 * md_oracle.c -- FP64 CPU restatement of the MD_Brushes force-and-step path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product never does.
 *
 * Each function restates one routine of wangshaoyun/MD_Brushes (cited
 * file:line).  The reference's state lives in Fortran module variables; here
 * it lives in orc_sys.  Arithmetic is written in the reference's order and
 * compiled with -ffp-contract=off.  Single-precision literals of the
 * reference (2**(1./6), 5**(1./6), 0.4, 1e-10, 1e15, 1.*i/Kmax) are kept in
 * single precision, as the Fortran compiler evaluates them.
 *
 * Pinned by tests/test_oracle.py against: closed forms, brute-force pair
 * enumeration, finite differences of the closed-form potentials, numpy.fft,
 * the B-spline recursion (compt_acc.f90:1667-1687), standard Ewald
 * (compt_acc.f90:1052-1135), the exact 2D-periodic Ewald sum (which the
 * vacuum-stretched 3D sum plus the slab correction must reach), Random123's
 * published Philox4x32-10 known-answer vectors and the integrator's
 * closed-form single-particle solution.
 */
#include "md_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Built with -fopenmp (liborc_fast.so, the all-cores CPU baseline) the hot
 * loops below run in parallel without changing any result: list loops are
 * split at particle boundaries (each particle's sum keeps its order), the
 * charge spreading is split by mesh planes (each point's sum keeps the charge
 * order), and the FFT, gather, noise and integrator loops are independent
 * per line or particle.  The checker build (liborc.so) is serial. */
#ifdef _OPENMP
#define ORC_NT omp_get_max_threads()
#define ORC_TID omp_get_thread_num()
#else
#define ORC_NT 1
#define ORC_TID 0
#endif

/* Thread t's share [b[t], b[t+1]) of a pair list grouped by its first index:
 * equal shares moved forward to the next change of that index. */
static void row_chunks(const int32_t *li, int64_t n, int nt, int64_t *b) {
  b[0] = 0;
  for (int t = 1; t < nt; t++) {
    int64_t k = n * t / nt;
    if (k < b[t - 1]) k = b[t - 1];
    while (k > 0 && k < n && li[k] == li[k - 1]) k++;
    b[t] = k;
  }
  b[nt] = n;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

#define PI 3.141592653589793
#define MAXO 12

struct orc_sys {
  md_params p;
  int NN, Nq, N_bond, N_anchor;
  double *pos, *vel, *acc, *acc_c;
  int32_t *charge;      /* 0-based particle index of charge m          */
  int32_t *fene;        /* fene[b], fene[N_bond+b], 0-based             */
  int32_t *anchor;      /* 0-based                                      */
  int32_t *l1i, *l1j;   /* LJ list (reference order)                    */
  int64_t npair1, cap1;
  int32_t *l2i, *l2j;   /* real-space list                              */
  int64_t npair2, cap2;
  double dr_max1, dr_max2;
  double bcof[3][MAXO * MAXO], dcof[3][MAXO * MAXO];
  double *BC;           /* ng1*ng2*ng3, x fastest                        */
  double complex *Q, *U;
};

/* ---------------------------------------------------------------- params */

/* data_operation (in_output.f90:217-288), read_force_parameters
 * (compt_acc.f90:159-162), Initialize_ewald_parameters
 * (compt_acc.f90:375-420). */
static void orc_ewald_rule(md_params *p, double tol, double tau_rf) {
  double Lx = p->Lx, Ly = p->Ly, Lz = p->Lz;
  int K1, K2, K3;
  double alpha = pow(tau_rf * PI * PI * PI * p->Nq / ((Lx * Ly * Lz) * (Lx * Ly * Lz)),
                     1.0 / 6);
  double rcc = tol / alpha;
  double rvc = rcc + p->rsk;
  if (p->ovr_alpha > 0) {
    alpha = p->ovr_alpha;
    rcc = p->ovr_rcc > 0 ? p->ovr_rcc : tol / alpha;
    rvc = rcc + p->rsk;
  }
  if ((int)(Lx / rvc) * (int)(Ly / rvc) * (int)(Lz / rvc) > 27) {
    K1 = (int)ceil(tol * Lx * alpha / PI);
    K2 = (int)ceil(tol * Ly * alpha / PI);
    K3 = (int)ceil(tol * Lz * p->Zfac * alpha / PI);
    p->real_verlet = 1;
  } else {
    rcc = (Lx > Ly) ? Ly / 2 : Lx / 2;
    K1 = (int)ceil(tol * tol / PI * Lx / rcc);
    K2 = (int)ceil(tol * tol / PI * Ly / rcc);
    K3 = (int)ceil(tol * tol / PI * Lz * p->Zfac / rcc);
    alpha = tol / rcc;
    rvc = rcc + p->rsk;
    p->real_verlet = 0;
  }
  K1 = (int)ceilf((float)K1 * 3.0f);
  K2 = (int)ceilf((float)K2 * 3.0f);
  K3 = (int)ceilf((float)K3 * 3.0f);
  p->alpha = alpha;
  p->rcc = rcc;
  p->rvc = rvc;
  p->ng[0] = p->ovr_ng[0] > 0 ? p->ovr_ng[0] : K1;
  p->ng[1] = p->ovr_ng[1] > 0 ? p->ovr_ng[1] : K2;
  p->ng[2] = p->ovr_ng[2] > 0 ? p->ovr_ng[2] : K3;
  p->gdim[0] = Lx;
  p->gdim[1] = Ly;
  p->gdim[2] = Lz * p->Zfac;
}

static int nint_d(double x) { return (int)lround(x); }

int orc_derive(md_params *p) {
  int arm_q = 0, i;
  if (p->arm < 0 || p->arm > MD_MAX_ARM) return MD_EINVAL;
  if (p->Nga != 0)
    for (i = 0; i < p->arm; i++) arm_q += p->charged_arm[i];
  p->Nta = (p->Nma * p->arm + 1) * p->Nga;
  p->Ntl = p->Nml * p->Ngl;
  p->Npe = p->Nta + p->Ntl;
  p->N_anchor = p->Nga + p->Ngl;
  if (fabs(p->qq) == 0) {
    p->Nq_PE = 0;
  } else if (p->man_l != 0 && p->man_s != 0) {
    p->Nq_PE = (int)floorf(1.0f * (p->Nml - 1) / p->man_l) * p->Ngl +
               (int)floorf(1.0f * p->Nma / p->man_s) * arm_q * p->Nga;
  } else {
    p->Nq_PE = 0;
  }
  if (fabs(p->qqi) == 0) {
    p->Nq_salt_ions = 0;
  } else {
    int charge_ions = nint_d(p->ion_ratio * p->Nq_PE * nint_d(fabs(p->qq)));
    p->Nq_salt_ions = charge_ions / nint_d(fabs(p->qqi));
  }
  p->Nq = p->Nq_PE * (nint_d(fabs(p->qq)) + 1) +
          p->Nq_salt_ions * (nint_d(fabs(p->qqi)) + 1);
  p->NN = p->Npe + p->Nq_PE * nint_d(fabs(p->qq)) +
          p->Nq_salt_ions * (nint_d(fabs(p->qqi)) + 1);
  p->N_bond = p->arm * p->Nma * p->Nga + (p->Nml - 1) * p->Ngl;
  if (p->N_anchor <= 0 || p->sigmag <= 0) return MD_EINVAL;
  p->Lx = sqrt(p->N_anchor / p->sigmag);
  p->Ly = p->Lx;
  p->Zfac = (p->Lz + p->Lx * p->Z_empty) / p->Lz;
  /* compt_acc.f90:159-162: 2**(1./6) is a default-real constant */
  p->sigma_sp = (p->sigma + p->sigma_s) / 2;
  {
    float c6 = (float)pow(2.0, (double)(1.0f / 6.0f));
    p->rcl_s = (double)c6 * p->sigma_s;
    p->rcl_sp = (double)c6 * p->sigma_sp;
  }
  /* the LJ list radius rcl_s + 1 (compt_acc.f90:162); a caller's skin
     (ovr_skin, an engine option) widens it the same way */
  p->rvl_eff = p->rcl_s + (p->ovr_skin > 0 ? 2 * p->ovr_skin : 1);
  if (p->qq != 0) {
    orc_ewald_rule(p, p->tol, p->tau_rf);
  } else {
    p->real_verlet = 0;
    p->alpha = p->rcc = p->rvc = 0;
    p->ng[0] = p->ng[1] = p->ng[2] = 0;
    p->gdim[0] = p->Lx;
    p->gdim[1] = p->Ly;
    p->gdim[2] = p->Lz * p->Zfac;
  }
  return MD_OK;
}

/* build_fene_list (compt_acc.f90:317-372) and build_charge
 * (compt_acc.f90:433-488); outputs 1-based like the reference. */
int orc_build_topology(const md_params *p, int32_t *charge, int32_t *fene,
                       int32_t *anchor) {
  int i, j, k, l;
  int arm = p->arm, Nma = p->Nma, Nga = p->Nga, Nml = p->Nml, Ngl = p->Ngl;
  int Nta = p->Nta, NB = p->N_bond, Npe = p->Npe, NN = p->NN;
  for (i = 1; i <= Nga; i++) anchor[i - 1] = (i - 1) * (arm * Nma + 1) + 1;
  for (i = 1; i <= Ngl; i++) anchor[i + Nga - 1] = Nta + (i - 1) * Nml + 1;
  l = 0;
  for (i = 1; i <= Nga; i++) {
    int base = (i - 1) * (arm * Nma + 1);
    for (k = 1; k <= Nma; k++) {
      l++;
      fene[l - 1] = base + k;
      fene[NB + l - 1] = base + k + 1;
    }
    for (j = 2; j <= arm; j++) {
      l++;
      fene[l - 1] = base + Nma + 1;
      fene[NB + l - 1] = base + (Nma * (j - 1) + 1) + 1;
      for (k = 1; k <= Nma - 1; k++) {
        l++;
        fene[l - 1] = base + (Nma * (j - 1) + 1) + k;
        fene[NB + l - 1] = base + (Nma * (j - 1) + 1) + k + 1;
      }
    }
  }
  for (i = 1; i <= Ngl; i++)
    for (j = 1; j <= Nml - 1; j++) {
      l++;
      fene[l - 1] = Nta + (i - 1) * Nml + j;
      fene[NB + l - 1] = Nta + (i - 1) * Nml + j + 1;
    }
  if (charge == NULL || p->qq == 0) return MD_OK;
  l = 0;
  for (i = 1; i <= Nga; i++) {
    int base = (i - 1) * (arm * Nma + 1);
    if (p->charged_arm[0] != 0)
      for (k = 2; k <= Nma + 1; k++)
        if ((k - 1) % p->man_s == 0) charge[l++] = base + k;
    for (j = 2; j <= arm; j++)
      if (p->charged_arm[j - 1] != 0)
        for (k = 1; k <= Nma; k++)
          if (k % p->man_s == 0) charge[l++] = base + ((j - 1) * Nma + 1) + k;
  }
  for (i = 1; i <= Ngl; i++)
    for (j = 2; j <= Nml; j++)
      if ((j - 1) % p->man_l == 0) charge[l++] = Nta + (i - 1) * Nml + j;
  for (i = Npe + 1; i <= NN; i++) charge[l++] = i;
  return (l == p->Nq) ? MD_OK : MD_EINVAL;
}

/* ------------------------------------------------------------- B-splines */

static double factorial(int n) { /* compt_acc.f90:1652-1664 */
  double f = 1.0;
  for (int i = n; i >= 1; i--) f *= i;
  return f;
}

static int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

/* bspln_coeffs (compt_acc.f90:1495-1554), Milovanovic's recurrence for the
 * piecewise polynomial coefficients of the cardinal B-spline of order ord.
 * cof[(k-1)*MAXO + (j-1)] = bspln_cof(k,j): piece k (x in [k-1,k]),
 * coefficient of x^(ord-j). */
void orc_bspln_coeffs(int ord, double *cof, double *dcof) {
  int D = ord + 2;
#define A(i, m, k) a[((i) * D + (m)) * D + (k)]
  double *a = calloc((size_t)D * D * D, sizeof(double));
  int m, k, i, ii, g;
  A(1, 1, 2) = 1;
  for (m = 2; m <= ord; m++) {
    A(m, m, 1) = 0;
    A(m - 1, m, 1) = 0;
    g = (int)floor(m / 2.0) - 1;
    for (k = 0; k <= g; k++) {
      A(m, m, k + 2) = 1.0 / (m - 1) / (m - k) *
                       (1.0 * m * A(m - 1, m - 1, k + 2) -
                        1.0 * k * (m - 1) * A(m, m, k + 1) - 1.0 * A(m - 1, m, k + 1));
      A(m, m, m - k + 1) = pow(-1.0, m - 1) * A(m, m, k + 2);
      for (i = m - 2; i >= 0; i--) {
        A(i + 1, m, k + 2) = 1.0 * m / (i + 1 - m) *
                             (1.0 * (i + 1) * A(i + 2, m, k + 2) - 1.0 * A(i + 1, m - 1, k + 2));
        for (ii = 0; ii <= m - i - 1; ii++)
          A(i + 1, m, m - k + 1) = 1.0 * A(i + 1, m, m - k + 1) +
                                   factorial(ii + i) / factorial(ii) *
                                       A(i + ii + 1, m, k + 2) * ipow(m, ii);
        A(i + 1, m, m - k + 1) = 1.0 * A(i + 1, m, m - k + 1) * ipow(-1, i) / factorial(i);
      }
    }
    if (m % 2 != 0) {
      A(m, m, g + 3) = 1.0 / (m - 1) / (m - (g + 1)) *
                       (1.0 * m * A(m - 1, m - 1, g + 3) -
                        1.0 * (g + 1) * (m - 1) * A(m, m, g + 2) - 1.0 * A(m - 1, m, g + 2));
      for (i = m - 2; i >= 0; i--)
        A(i + 1, m, g + 3) = 1.0 * m / (i + 1 - m) *
                             (1.0 * (i + 1) * A(i + 2, m, g + 3) - 1.0 * A(i + 1, m - 1, g + 3));
    }
  }
  for (i = 0; i < MAXO * MAXO; i++) cof[i] = dcof[i] = 0;
  for (i = 0; i <= ord - 1; i++)
    for (k = 0; k <= ord - 1; k++) cof[k * MAXO + i] = A(ord - i, ord, k + 2);
  for (i = 1; i <= ord; i++)
    for (int j = 1; j <= ord; j++)
      dcof[(i - 1) * MAXO + (j - 1)] = cof[(i - 1) * MAXO + (j - 1)] * (ord - j);
#undef A
  free(a);
}

/* bspln (compt_acc.f90:1667-1687): Cox-de Boor recursion. */
double orc_bspln(double x, int ord) {
  if (ord > 2)
    return (x / (ord - 1)) * orc_bspln(x, ord - 1) +
           ((ord - x) / (ord - 1)) * orc_bspln(x - 1.0, ord - 1);
  if (ord == 2) return (x >= 0.0 && x <= 2.0) ? 1.0 - fabs(x - 1.0) : 0.0;
  return 0.0;
}

/* Weights of splcof (compt_acc.f90:1265-1302) for n fractional offsets w in
 * (0,1]: sp[t*ord + (j-1)] = SP(t,j), dsp likewise; no charge factor. */
void orc_splcof(int ord, const double *w_in, int n, const double *cof,
                const double *dcof, double *sp, double *dsp) {
  for (int t = 0; t < n; t++) {
    double *S = sp + (size_t)t * ord, *dS = dsp + (size_t)t * ord;
    double ww = w_in[t];
    if (ord == 4) { /* closed form, compt_acc.f90:1272-1287 */
      double s1, s2, s3;
      s3 = 0.5 * ww * ww;
      s1 = 0.5 * (1.0 - ww) * (1.0 - ww);
      s2 = 1.0 - s1 - s3;
      dS[0] = -s1;
      dS[1] = s1 - s2;
      dS[2] = s2 - s3;
      dS[3] = s3;
      S[3] = ww * s3 / 3;
      S[2] = ((ww + 1.0) * s2 + (3.0 - ww) * s3) / 3;
      S[0] = (1.0 - ww) * s1 / 3;
      S[1] = 1.0 - S[0] - S[2] - S[3];
    } else { /* polynomial pieces, compt_acc.f90:1288-1302 */
      for (int k = 1; k <= ord; k++) {
        const double *c = cof + (k - 1) * MAXO, *d = dcof + (k - 1) * MAXO;
        double s = c[0] * ww, ds = d[0] * ww;
        for (int j = 2; j <= ord - 2; j++) {
          s = ww * (s + c[j - 1]);
          ds = ww * (ds + d[j - 1]);
        }
        s = ww * (s + c[ord - 2]);
        s = s + c[ord - 1];
        ds = ds + d[ord - 2];
        S[ord - k] = s;
        dS[ord - k] = ds;
        ww = ww + 1;
      }
    }
  }
}

/* pmeOrthoTabBC (compt_acc.f90:1601-1649) for one axis: b2 = |b(m)|^2,
 * mv = folded wave number, ev = exp(-(pi/alpha)^2 mv^2) (no pivol). */
void orc_pme_bc_axis(int ng, int ord, double gdim, double alpha, double *b2,
                     double *mv, double *ev) {
  double tpS = PI / alpha;
  for (int i = 1; i <= ng; i++) {
    mv[i - 1] = (i - 1 <= ng / 2) ? (i - 1) / gdim : (i - 1 - ng) / gdim;
    ev[i - 1] = exp(-tpS * tpS * mv[i - 1] * mv[i - 1]);
  }
  for (int i = 0; i <= ng - 1; i++) {
    double complex bsp = 0;
    for (int j = 0; j <= ord - 2; j++)
      bsp = bsp + orc_bspln(j + 1.0, ord) *
                      (cos(2 * PI * i * j / ng) + I * sin(2 * PI * i * j / ng));
    if (cabs(bsp) < (double)1e-10f) bsp = (double)1e15f;
    double complex bv =
        (cos(2 * PI * i * (ord - 1) / ng) + I * sin(2 * PI * i * (ord - 1) / ng)) / bsp;
    b2[i] = creal(bv * conj(bv));
  }
}

/* ------------------------------------------------------------------- FFT */

/* Unnormalised DFT of one line by mixed-radix decimation in time: the p
 * sub-transforms of length n/p (p the smallest prime factor of n) over the
 * inputs r, r + p, ..., then X[k + m q] = sum_r w^(r (k + m q) N/n) Y_r[k].
 * w[k] = exp(sign 2 pi i k / N) is the table of the whole line's length N,
 * computed once per axis; tmp holds n points per level. */
static int smallest_factor(int n) {
  if (n % 2 == 0) return 2;
  for (int f = 3; f * f <= n; f += 2)
    if (n % f == 0) return f;
  return n;
}

static void fft_rec(const double complex *x, int xs, double complex *X, int n, int N,
                    const double complex *w, double complex *tmp) {
  if (n == 1) {
    X[0] = x[0];
    return;
  }
  int p = smallest_factor(n), m = n / p, wstep = N / n;
  for (int r = 0; r < p; r++) fft_rec(x + (size_t)r * xs, xs * p, X + (size_t)r * m, m, N, w, tmp + n);
  for (int q = 0; q < p; q++)
    for (int k = 0; k < m; k++) {
      double complex s = 0;
      for (int r = 0; r < p; r++) {
        long e = ((long)r * (k + (long)m * q)) % n;
        s += X[(size_t)r * m + k] * w[e * wstep];
      }
      tmp[k + (size_t)m * q] = s;
    }
  memcpy(X, tmp, sizeof(double complex) * n);
}

static double complex *twiddles(int n, int sign) {
  double complex *w = malloc(sizeof(double complex) * n);
  for (int k = 0; k < n; k++) {
    double ang = sign * 2 * PI * (double)k / n;
    w[k] = cos(ang) + I * sin(ang);
  }
  return w;
}

/* Unnormalised 3D DFT, FFTW sign convention (fftw3.f90:12-13: forward -1,
 * backward +1), x fastest: the transform of pmeOrthoConvBC
 * (compt_acc.f90:1382-1396). */
static void fft3d_c(double complex *a, int n1, int n2, int n3, int sign) {
  int nmax = n1 > n2 ? n1 : n2;
  nmax = nmax > n3 ? nmax : n3;
  double complex *w1 = twiddles(n1, sign), *w2 = twiddles(n2, sign), *w3 = twiddles(n3, sign);
#pragma omp parallel
  {
  double complex *line = malloc(sizeof(double complex) * nmax);
  double complex *out = malloc(sizeof(double complex) * nmax);
  /* scratch: n points per recursion level, at most 2n in all */
  double complex *tmp = malloc(sizeof(double complex) * 2 * nmax);
#pragma omp for
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++) {
      double complex *r = a + ((size_t)k * n2 + j) * n1;
      fft_rec(r, 1, out, n1, n1, w1, tmp);
      memcpy(r, out, sizeof(double complex) * n1);
    }
#pragma omp for
  for (int k = 0; k < n3; k++)
    for (int i = 0; i < n1; i++) {
      for (int j = 0; j < n2; j++) line[j] = a[((size_t)k * n2 + j) * n1 + i];
      fft_rec(line, 1, out, n2, n2, w2, tmp);
      for (int j = 0; j < n2; j++) a[((size_t)k * n2 + j) * n1 + i] = out[j];
    }
#pragma omp for
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n1; i++) {
      for (int k = 0; k < n3; k++) line[k] = a[((size_t)k * n2 + j) * n1 + i];
      fft_rec(line, 1, out, n3, n3, w3, tmp);
      for (int k = 0; k < n3; k++) a[((size_t)k * n2 + j) * n1 + i] = out[k];
    }
  free(line);
  free(out);
  free(tmp);
  }
  free(w1);
  free(w2);
  free(w3);
}

void orc_fft3d(int n1, int n2, int n3, double *re_im, int sign) {
  fft3d_c((double complex *)re_im, n1, n2, n3, sign);
}

/* ------------------------------------------------------------ system */

orc_sys *orc_create(const md_params *pin) {
  orc_sys *s = calloc(1, sizeof(orc_sys));
  s->p = *pin;
  md_params *p = &s->p;
  s->NN = p->NN;
  s->Nq = p->Nq;
  s->N_bond = p->N_bond;
  s->N_anchor = p->N_anchor;
  s->pos = calloc((size_t)s->NN * 4, sizeof(double));
  s->vel = calloc((size_t)s->NN * 3, sizeof(double));
  s->acc = calloc((size_t)s->NN * 3, sizeof(double));
  s->acc_c = calloc((size_t)s->NN * 3, sizeof(double));
  s->charge = calloc((size_t)(s->Nq > 0 ? s->Nq : 1), sizeof(int32_t));
  s->fene = calloc((size_t)(2 * s->N_bond + 2), sizeof(int32_t));
  s->anchor = calloc((size_t)(s->N_anchor + 1), sizeof(int32_t));
  orc_build_topology(p, s->Nq > 0 ? s->charge : NULL, s->fene, s->anchor);
  for (int m = 0; m < s->Nq; m++) s->charge[m]--;
  for (int b = 0; b < 2 * s->N_bond; b++) s->fene[b]--;
  for (int a = 0; a < s->N_anchor; a++) s->anchor[a]--;
  if (p->qq != 0) {
    /* bspln_coeffs + PME_BC (compt_acc.f90:127-130,1557-1598) */
    for (int h = 0; h < 3; h++) orc_bspln_coeffs(p->ordr[h], s->bcof[h], s->dcof[h]);
    int n1 = p->ng[0], n2 = p->ng[1], n3 = p->ng[2];
    size_t M = (size_t)n1 * n2 * n3;
    s->BC = calloc(M, sizeof(double));
    s->Q = calloc(M, sizeof(double complex));
    s->U = calloc(M, sizeof(double complex));
    double *b2[3], *mv[3], *ev[3];
    for (int h = 0; h < 3; h++) {
      b2[h] = malloc(sizeof(double) * p->ng[h]);
      mv[h] = malloc(sizeof(double) * p->ng[h]);
      ev[h] = malloc(sizeof(double) * p->ng[h]);
      orc_pme_bc_axis(p->ng[h], p->ordr[h], p->gdim[h], p->alpha, b2[h], mv[h], ev[h]);
    }
    double pivol = (double)n1 * n2 * n3 / (PI * p->gdim[0] * p->gdim[1] * p->gdim[2]);
    for (int i = 0; i < n1; i++) ev[0][i] *= pivol;
    for (int i = 1; i <= n1; i++) {
      double dx = mv[0][i - 1] * mv[0][i - 1];
      for (int j = 1; j <= n2; j++) {
        double dxy = dx + mv[1][j - 1] * mv[1][j - 1];
        for (int k = 1; k <= n3; k++) {
          double dxyz = dxy + mv[2][k - 1] * mv[2][k - 1];
          if (i + j + k > 3)
            s->BC[((size_t)(k - 1) * n2 + (j - 1)) * n1 + (i - 1)] =
                b2[0][i - 1] * b2[1][j - 1] * b2[2][k - 1] * ev[0][i - 1] *
                ev[1][j - 1] * ev[2][k - 1] / dxyz;
        }
      }
    }
    for (int h = 0; h < 3; h++) {
      free(b2[h]);
      free(mv[h]);
      free(ev[h]);
    }
  }
  return s;
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
  free(s->pos); free(s->vel); free(s->acc); free(s->acc_c);
  free(s->charge); free(s->fene); free(s->anchor);
  free(s->l1i); free(s->l1j); free(s->l2i); free(s->l2j);
  free(s->BC); free(s->Q); free(s->U);
  free(s);
}

void orc_set_state(orc_sys *s, const double *pos, const double *vel,
                   const double *acc) {
  if (pos) memcpy(s->pos, pos, sizeof(double) * 4 * s->NN);
  if (vel) memcpy(s->vel, vel, sizeof(double) * 3 * s->NN);
  if (acc) memcpy(s->acc, acc, sizeof(double) * 3 * s->NN);
}

void orc_get_state(orc_sys *s, double *pos, double *vel, double *acc,
                   double *acc_c, double *dr_max) {
  if (pos) memcpy(pos, s->pos, sizeof(double) * 4 * s->NN);
  if (vel) memcpy(vel, s->vel, sizeof(double) * 3 * s->NN);
  if (acc) memcpy(acc, s->acc, sizeof(double) * 3 * s->NN);
  if (acc_c) memcpy(acc_c, s->acc_c, sizeof(double) * 3 * s->NN);
  if (dr_max) {
    dr_max[0] = s->dr_max1;
    dr_max[1] = s->dr_max2;
  }
}

void orc_set_drmax(orc_sys *s, double d1, double d2) {
  s->dr_max1 = d1;
  s->dr_max2 = d2;
}

#define X(i) s->pos[(i)]
#define Y(i) s->pos[(size_t)s->NN + (i)]
#define Z(i) s->pos[2 * (size_t)s->NN + (i)]
#define QC(i) s->pos[3 * (size_t)s->NN + (i)]

/* rij_and_rr (glob_var.f90:72-105): x,y minimum image, z open. */
static inline double rij_rr(const orc_sys *s, int i, int j, double r[3]) {
  double Lx = s->p.Lx, Ly = s->p.Ly;
  r[0] = X(i) - X(j);
  r[1] = Y(i) - Y(j);
  r[2] = Z(i) - Z(j);
  if (r[0] > Lx / 2) r[0] = r[0] - Lx;
  else if (r[0] <= -Lx / 2) r[0] = r[0] + Lx;
  if (r[1] > Ly / 2) r[1] = r[1] - Ly;
  else if (r[1] <= -Ly / 2) r[1] = r[1] + Ly;
  return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

/* ------------------------------------------------------- Verlet lists */

/* Shared body of lj_verlet_list (compt_acc.f90:185-314) and
 * real_verlet_list (compt_acc.f90:491-626): linked cells of edge >= rv,
 * 27-cell stencil periodic in x,y and open in z, full (i,j)+(j,i) list,
 * count pass then fill pass.  idx maps list slot -> particle (NULL =
 * identity). */
static int64_t build_list(orc_sys *s, int n, const int32_t *idx, double rv,
                          int32_t **li, int32_t **lj, int64_t *cap) {
  double Lx = s->p.Lx, Ly = s->p.Ly, Lz = s->p.Lz;
  int nc1 = (int)(Lx / rv), nc2 = (int)(Ly / rv), nc3 = (int)(Lz / rv);
  if (nc1 < 1 || nc2 < 1 || nc3 < 1) return -1;
  double rc1 = Lx / nc1, rc2 = Ly / nc2, rc3 = Lz / nc3;
  int *hoc = calloc((size_t)nc1 * nc2 * nc3, sizeof(int));
  int *cl = calloc((size_t)n + 1, sizeof(int));
  int *ci = malloc(sizeof(int) * 3 * (size_t)(n + 1));
  for (int m = 1; m <= n; m++) {
    int i = idx ? idx[m - 1] : m - 1;
    int a = (int)((X(i) + Lx / 2) / rc1), b = (int)((Y(i) + Ly / 2) / rc2),
        c = (int)(Z(i) / rc3);
    if (a < 0 || a >= nc1 || b < 0 || b >= nc2 || c < 0 || c >= nc3) {
      free(hoc); free(cl); free(ci);
      return -2;
    }
    ci[3 * m] = a; ci[3 * m + 1] = b; ci[3 * m + 2] = c;
    size_t h = ((size_t)c * nc2 + b) * nc1 + a;
    cl[m] = hoc[h];
    hoc[h] = m;
  }
  /* two passes over the targets m: the hits of each (count), then the
   * pairs at each target's offset (fill); targets in order, so the list is
   * the reference's whether or not the passes run in parallel */
  int64_t *off = calloc((size_t)n + 2, sizeof(int64_t));
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int m = 1; m <= n; m++) off[m + 1] += off[m];
      if (off[n + 1] > *cap) {
        free(*li); free(*lj);
        *cap = off[n + 1] + 1;
        *li = malloc(sizeof(int32_t) * *cap);
        *lj = malloc(sizeof(int32_t) * *cap);
      }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int m = 1; m <= n; m++) {
      int i = idx ? idx[m - 1] : m - 1;
      int a = ci[3 * m], b = ci[3 * m + 1], c = ci[3 * m + 2];
      int64_t k = pass == 1 ? off[m] : 0;
      for (int l = -1; l <= 1; l++) {
        int pp = a + l >= nc1 ? a + l - nc1 : (a + l < 0 ? a + l + nc1 : a + l);
        for (int mm = -1; mm <= 1; mm++) {
          int qq = b + mm >= nc2 ? b + mm - nc2 : (b + mm < 0 ? b + mm + nc2 : b + mm);
          for (int nn = -1; nn <= 1; nn++) {
            if (c + nn >= nc3 || c + nn < 0) continue;
            int r = c + nn;
            int t = hoc[((size_t)r * nc2 + qq) * nc1 + pp];
            while (t != 0) {
              int j = idx ? idx[t - 1] : t - 1;
              double rij[3];
              double rsqr = rij_rr(s, i, j, rij);
              if (i != j && rsqr < rv * rv) {
                if (pass == 1) {
                  (*li)[k] = i;
                  (*lj)[k] = j;
                }
                k++;
              }
              t = cl[t];
            }
          }
        }
      }
      if (pass == 0) off[m + 1] = k;
    }
  }
  int64_t k = off[n + 1];
  free(off);
  free(hoc); free(cl); free(ci);
  return k;
}

int64_t orc_lj_verlet_list(orc_sys *s) {
  int64_t k = build_list(s, s->NN, NULL, s->p.rvl_eff, &s->l1i, &s->l1j, &s->cap1);
  if (k >= 0) s->npair1 = k;
  return k;
}

int64_t orc_real_verlet_list(orc_sys *s) {
  int64_t k = build_list(s, s->Nq, s->charge, s->p.rvc, &s->l2i, &s->l2j, &s->cap2);
  if (k >= 0) s->npair2 = k;
  return k;
}

/* pairs(cap,2) column-major, 1-based, in reference order. */
int64_t orc_get_list(orc_sys *s, int which, int32_t *pairs, int64_t cap) {
  int64_t n = which == 0 ? s->npair1 : s->npair2;
  int32_t *a = which == 0 ? s->l1i : s->l2i, *b = which == 0 ? s->l1j : s->l2j;
  for (int64_t k = 0; k < n && k < cap; k++) {
    pairs[k] = a[k] + 1;
    pairs[cap + k] = b[k] + 1;
  }
  return n;
}

/* ------------------------------------------------------ short range */

/* lj_force (compt_acc.f90:895-977): WCA with three size classes chosen by
 * |q_i|+|q_j|, full list so only i is updated, then the wall force.
 * f(NN,3) is overwritten. */
int orc_lj_force(orc_sys *s, double *f) {
  md_params *p = &s->p;
  int NN = s->NN;
  double Lx = p->Lx, Ly = p->Ly, Lz = p->Lz;
  double hLx = Lx / 2, nhLx = -Lx / 2, hLy = Ly / 2, nhLy = -Ly / 2;
  double rcl2 = p->rcl * p->rcl, rcl3 = p->rcl_s * p->rcl_s, rcl4 = p->rcl_sp * p->rcl_sp;
  double sigma2 = p->sigma_s * p->sigma_s, sigma3 = p->sigma_sp * p->sigma_sp;
  memset(f, 0, sizeof(double) * 3 * NN);
  int64_t *rb = malloc(sizeof(int64_t) * (ORC_NT + 1));
  row_chunks(s->l1i, s->npair1, ORC_NT, rb);
#pragma omp parallel
  {
  for (int64_t k = rb[ORC_TID]; k < rb[ORC_TID + 1]; k++) {
    int i = s->l1i[k], j = s->l1j[k];
    double rij[3] = {X(i) - X(j), Y(i) - Y(j), Z(i) - Z(j)};
    int m = (int)(fabs(QC(i)) + fabs(QC(j)));
    if (rij[0] > hLx) rij[0] = rij[0] - Lx;
    else if (rij[0] <= nhLx) rij[0] = rij[0] + Lx;
    if (rij[1] > hLy) rij[1] = rij[1] - Ly;
    else if (rij[1] <= nhLy) rij[1] = rij[1] + Ly;
    double rsqr = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    double inv_r2, inv_r6;
    int hit = 0;
    if (m == 6) {
      if (rsqr < rcl3) { inv_r2 = sigma2 / rsqr; hit = 1; }
    } else if (m < 3) {
      if (rsqr < rcl2) { inv_r2 = 1 / rsqr; hit = 1; }
    } else {
      if (rsqr < rcl4) { inv_r2 = sigma3 / rsqr; hit = 1; }
    }
    if (hit) {
      inv_r6 = inv_r2 * inv_r2 * inv_r2;
      for (int d = 0; d < 3; d++)
        f[(size_t)d * NN + i] = f[(size_t)d * NN + i] + inv_r6 * (inv_r6 - 0.5) * rij[d] / rsqr;
    }
  }
  }
  free(rb);
#pragma omp parallel for
  for (int i = 0; i < 3 * NN; i++) f[i] = f[i] * 48;
  /* wall (compt_acc.f90:955-975); 0.4 and 5**(1./6) are default-real */
  const double c04 = (double)0.4f;
  const double c5 = (double)(float)pow(5.0, (double)(1.0f / 6.0f));
  double *fz = f + 2 * (size_t)NN;
  for (int i = 0; i < NN; i++) {
    double z = Z(i), dl;
    if (fabs(z) < 0.01 || fabs(z - Lz) < 0.01) continue;
    if (QC(i) != 3) {
      if (z < p->rcl / c5) {
        dl = z;
        fz[i] = fz[i] + c04 * (3 / pow(dl, 11) - 1 / pow(dl, 5)) * dl;
      } else if ((Lz - z) < p->rcl / c5) {
        dl = Lz - z;
        fz[i] = fz[i] - c04 * (3 / pow(dl, 11) - 1 / pow(dl, 5)) * dl;
      }
    } else {
      double ss = p->sigma_s;
      if (z < p->rcl_s / c5) {
        dl = z;
        fz[i] = fz[i] + c04 * (3 * pow(ss / dl, 11) - 1 * pow(ss / dl, 5)) * dl / ss / ss;
      } else if ((Lz - z) < p->rcl_s / c5) {
        dl = Lz - z;
        fz[i] = fz[i] - c04 * (3 * pow(ss / dl, 11) - 1 * pow(ss / dl, 5)) * dl / ss / ss;
      }
    }
  }
  return MD_OK;
}

/* fene_force (compt_acc.f90:865-892).  Returns MD_EBOND where the reference
 * prints and stops; f(NN,3) is overwritten. */
int orc_fene_force(orc_sys *s, double *f) {
  int NN = s->NN;
  memset(f, 0, sizeof(double) * 3 * NN);
  for (int b = 0; b < s->N_bond; b++) {
    int j = s->fene[b], k = s->fene[s->N_bond + b];
    double r[3];
    double rsqr = rij_rr(s, j, k, r);
    if (rsqr > s->p.R0_2) return MD_EBOND;
    for (int d = 0; d < 3; d++) {
      double ff = -s->p.kFENE * s->p.R0_2 / (s->p.R0_2 - rsqr) * r[d];
      f[(size_t)d * NN + j] += ff;
      f[(size_t)d * NN + k] -= ff;
    }
  }
  return MD_OK;
}

/* ------------------------------------------------------------ Coulomb */

/* real_space (compt_acc.f90:980-1049): erfc-screened pair force over the
 * real-space list (or all pairs when real_verlet == 0), then the
 * Yeh-Berkowitz slab correction.  Adds into f(NN,3). */
void orc_real_space(orc_sys *s, double *f) {
  md_params *p = &s->p;
  int NN = s->NN;
  double Mz = 0, c1 = p->lb / p->Beta, c2 = 2 * p->alpha / sqrt(PI);
  double alpha2 = p->alpha * p->alpha, rcc2 = p->rcc * p->rcc;
  double Lx = p->Lx, Ly = p->Ly;
  if (p->real_verlet == 0) {
    for (int m = 0; m < s->Nq - 1; m++) {
      int i = s->charge[m];
      for (int n = m + 1; n < s->Nq; n++) {
        int j = s->charge[n];
        double rij[3];
        double rsqr = rij_rr(s, i, j, rij);
        if (rsqr < rcc2) {
          double rr = sqrt(rsqr);
          for (int d = 0; d < 3; d++) {
            double ff = QC(i) * QC(j) / rsqr *
                        (erfc(p->alpha * rr) / rr + c2 * exp(-alpha2 * rsqr)) * rij[d] * c1;
            f[(size_t)d * NN + i] += ff;
            f[(size_t)d * NN + j] -= ff;
          }
        }
      }
    }
  } else {
    int64_t *rb = malloc(sizeof(int64_t) * (ORC_NT + 1));
    row_chunks(s->l2i, s->npair2, ORC_NT, rb);
#pragma omp parallel
    {
    for (int64_t k = rb[ORC_TID]; k < rb[ORC_TID + 1]; k++) {
      int i = s->l2i[k], j = s->l2j[k];
      double rij[3] = {X(i) - X(j), Y(i) - Y(j), Z(i) - Z(j)};
      if (rij[0] > Lx / 2) rij[0] = rij[0] - Lx;
      else if (rij[0] <= -Lx / 2) rij[0] = rij[0] + Lx;
      if (rij[1] > Ly / 2) rij[1] = rij[1] - Ly;
      else if (rij[1] <= -Ly / 2) rij[1] = rij[1] + Ly;
      double rsqr = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
      if (rsqr < rcc2) {
        double rr = sqrt(rsqr);
        for (int d = 0; d < 3; d++) {
          double ff = QC(i) * QC(j) / rsqr *
                      (erfc(p->alpha * rr) / rr + c2 * exp(-alpha2 * rsqr)) * rij[d] * c1;
          f[(size_t)d * NN + i] += ff;
        }
      }
    }
    }
    free(rb);
  }
  for (int m = 0; m < s->Nq; m++) {
    int i = s->charge[m];
    Mz = Mz + QC(i) * Z(i) * p->lb / p->Beta;
  }
  for (int m = 0; m < s->Nq; m++) {
    int i = s->charge[m];
    f[2 * (size_t)NN + i] = f[2 * (size_t)NN + i] -
                            4 * PI / Lx / Ly / p->Lz / p->Zfac * QC(i) * Mz;
  }
}

/* SPME_Ewald (compt_acc.f90:1213-1246): splcof, MapCharges, pmeOrthoConvBC,
 * IFrc, ZeroForce, scaled by lb/Beta.  Adds into f(NN,3). */
void orc_spme(orc_sys *s, double *f) {
  md_params *p = &s->p;
  int Nq = s->Nq, NN = s->NN;
  int n1 = p->ng[0], n2 = p->ng[1], n3 = p->ng[2];
  int o[3] = {p->ordr[0], p->ordr[1], p->ordr[2]};
  size_t M = (size_t)n1 * n2 * n3;
  double *SP[3], *dSP[3];
  int *map[3];
  for (int h = 0; h < 3; h++) {
    SP[h] = malloc(sizeof(double) * Nq * o[h]);
    dSP[h] = malloc(sizeof(double) * Nq * o[h]);
    map[h] = malloc(sizeof(int) * Nq * o[h]);
  }
  /* splcof (compt_acc.f90:1250-1321) */
  for (int h = 0; h < 3; h++) {
#pragma omp parallel for
    for (int m = 0; m < Nq; m++) {
      int i = s->charge[m];
      double uu = p->ng[h] / p->gdim[h] * s->pos[(size_t)h * NN + i];
      int uuc = (int)ceil(uu);
      double ww = uu - uuc + 1;
      orc_splcof(o[h], &ww, 1, s->bcof[h], s->dcof[h], SP[h] + (size_t)m * o[h],
                 dSP[h] + (size_t)m * o[h]);
      for (int t = 0; t < o[h]; t++) {
        int q = uuc - o[h] + t;
        q = q - p->ng[h] * (int)floorf(1.0f * q / p->ng[h]);
        map[h][(size_t)m * o[h] + t] = q;
      }
      if (h == 0)
        for (int t = 0; t < o[h]; t++) {
          SP[h][(size_t)m * o[h] + t] *= QC(i);
          dSP[h][(size_t)m * o[h] + t] *= QC(i);
        }
    }
  }
  /* MapCharges (compt_acc.f90:1324-1365) */
  memset(s->Q, 0, sizeof(double complex) * M);
#pragma omp parallel
  {
  const int z0 = (int)((long)n3 * ORC_TID / ORC_NT), z1 = (int)((long)n3 * (ORC_TID + 1) / ORC_NT);
  for (int m = 0; m < Nq; m++)
    for (int c = 0; c < o[2]; c++) {
      const int zc = map[2][m * o[2] + c];
      if (zc < z0 || zc >= z1) continue;  /* this thread's planes */
      for (int b = 0; b < o[1]; b++)
        for (int a = 0; a < o[0]; a++) {
          size_t g = ((size_t)zc * n2 + map[1][m * o[1] + b]) * n1 + map[0][m * o[0] + a];
          s->Q[g] += SP[2][m * o[2] + c] * (SP[0][m * o[0] + a] * SP[1][m * o[1] + b]);
        }
    }
  }
  /* pmeOrthoConvBC (compt_acc.f90:1368-1398) */
  memcpy(s->U, s->Q, sizeof(double complex) * M);
  fft3d_c(s->U, n1, n2, n3, -1);
  s->U[0] = 0;
#pragma omp parallel for
  for (size_t g = 0; g < M; g++) s->U[g] = s->U[g] * s->BC[g];
  fft3d_c(s->U, n1, n2, n3, +1);
  /* IFrc (compt_acc.f90:1402-1467) */
  double *af = calloc((size_t)Nq * 3, sizeof(double));
  double nn = (double)(n1 * n2 * n3);
#pragma omp parallel for
  for (int m = 0; m < Nq; m++) {
    double complex f1 = 0, f2 = 0, f3 = 0;
    for (int c = 0; c < o[2]; c++)
      for (int b = 0; b < o[1]; b++)
        for (int a = 0; a < o[0]; a++) {
          size_t g = ((size_t)map[2][m * o[2] + c] * n2 + map[1][m * o[1] + b]) * n1 +
                     map[0][m * o[0] + a];
          double Bx = SP[0][m * o[0] + a], By = SP[1][m * o[1] + b], Bz = SP[2][m * o[2] + c];
          double dBx = dSP[0][m * o[0] + a], dBy = dSP[1][m * o[1] + b],
                 dBz = dSP[2][m * o[2] + c];
          f1 += (Bz * (dBx * By)) * s->U[g];
          f2 += (Bz * (Bx * dBy)) * s->U[g];
          f3 += (dBz * (Bx * By)) * s->U[g];
        }
    af[m] = creal(-f1) * n1 / p->gdim[0] / nn;
    af[Nq + m] = creal(-f2) * n2 / p->gdim[1] / nn;
    af[2 * Nq + m] = creal(-f3) * n3 / p->gdim[2] / nn;
  }
  /* ZeroForce (compt_acc.f90:1470-1492) */
  for (int d = 0; d < 3; d++) {
    double mean = 0;
    for (int m = 0; m < Nq; m++) mean = mean + af[d * Nq + m];
    mean = mean / Nq;
    for (int m = 0; m < Nq; m++) af[d * Nq + m] = af[d * Nq + m] - mean;
  }
  /* compt_acc.f90:1244 */
  for (int m = 0; m < Nq; m++)
    for (int d = 0; d < 3; d++)
      f[(size_t)d * NN + s->charge[m]] += af[d * Nq + m] * p->lb / p->Beta;
  free(af);
  for (int h = 0; h < 3; h++) {
    free(SP[h]); free(dSP[h]); free(map[h]);
  }
}

/* fourier_function + Standard_Ewald (compt_acc.f90:629-671,1052-1135):
 * k-space Ewald sum over the ellipsoid kcut <= 1 (kcut in default real),
 * half space k3 >= 0 with weight 2.  Adds into f(NN,3). */
void orc_standard_ewald(orc_sys *s, double alpha, int K1, int K2, int K3,
                        double *f) {
  md_params *p = &s->p;
  int Nq = s->Nq, NN = s->NN;
  double Lx = p->Lx, Ly = p->Ly, Lz = p->Lz, Ze = p->Zfac;
  double alpha2 = alpha * alpha;
  double c1 = 2 * PI / Lx, c2 = 2 * PI / Ly, c3 = 2 * PI / Lz / Ze;
  size_t w1 = 2 * K1 + 1, w2 = 2 * K2 + 1, w3 = K3 + 1;
  double complex *ex = malloc(sizeof(double complex) * Nq * w1);
  double complex *ey = malloc(sizeof(double complex) * Nq * w2);
  double complex *ez = malloc(sizeof(double complex) * Nq * w3);
  double complex *eikr = malloc(sizeof(double complex) * Nq);
  double *zq = malloc(sizeof(double) * Nq);
  double *af = calloc((size_t)Nq * 3, sizeof(double));
#define EX(m, k) ex[(size_t)(m) * w1 + (k) + K1]
#define EY(m, k) ey[(size_t)(m) * w2 + (k) + K2]
#define EZ(m, k) ez[(size_t)(m) * w3 + (k)]
  for (int m = 0; m < Nq; m++) {
    int i = s->charge[m];
    zq[m] = QC(i);
    EX(m, 0) = 1; EY(m, 0) = 1; EZ(m, 0) = 1;
    EX(m, 1) = cos(c1 * X(i)) + I * sin(c1 * X(i));
    EY(m, 1) = cos(c2 * Y(i)) + I * sin(c2 * Y(i));
    EZ(m, 1) = cos(c3 * Z(i)) + I * sin(c3 * Z(i));
    EX(m, -1) = conj(EX(m, 1));
    EY(m, -1) = conj(EY(m, 1));
  }
  for (int q = 2; q <= K1; q++)
    for (int m = 0; m < Nq; m++) { EX(m, q) = EX(m, q - 1) * EX(m, 1); EX(m, -q) = conj(EX(m, q)); }
  for (int q = 2; q <= K2; q++)
    for (int m = 0; m < Nq; m++) { EY(m, q) = EY(m, q - 1) * EY(m, 1); EY(m, -q) = conj(EY(m, q)); }
  for (int q = 2; q <= K3; q++)
    for (int m = 0; m < Nq; m++) EZ(m, q) = EZ(m, q - 1) * EZ(m, 1);
  for (int r = 0; r <= K3; r++) {
    double factor = (r == 0) ? 1 : 2;
    for (int pp = -K1; pp <= K1; pp++)
      for (int q = -K2; q <= K2; q++) {
        float a = (1.0f * pp / K1), b = (1.0f * q / K2), c = (1.0f * r / K3);
        float kcut = a * a + b * b + c * c;
        if (kcut > 1 || kcut == 0) continue;
        double kk[3] = {pp * c1, q * c2, r * c3};
        double ksqr = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
        double ek = factor * 4 * PI / (Lx * Ly * Lz * Ze) * exp(-ksqr / 4 / alpha2) / ksqr *
                    p->lb / p->Beta;
        double complex sum = 0;
        for (int m = 0; m < Nq; m++) {
          eikr[m] = zq[m] * EX(m, pp) * EY(m, q) * EZ(m, r);
          sum = sum + eikr[m];
        }
        for (int m = 0; m < Nq; m++)
          for (int d = 0; d < 3; d++)
            af[d * Nq + m] += cimag(eikr[m] * (conj(sum) * kk[d])) * ek;
      }
  }
  for (int m = 0; m < Nq; m++)
    for (int d = 0; d < 3; d++) f[(size_t)d * NN + s->charge[m]] += af[d * Nq + m];
#undef EX
#undef EY
#undef EZ
  free(ex); free(ey); free(ez); free(eikr); free(zq); free(af);
}

/* error_analysis (compt_acc.f90:729-812).  Estimates of the production
 * parameters (Kolafa & Perram 1992, compt_acc.f90:746-751; the exponents
 * (1./3) are default real, and (Kmax1*Kmax2*Kmax3)**(1./3) is evaluated in
 * default real); then the reference Coulomb force: parameters of the rule
 * at tol = 5, tau_rf = 10 without overrides (compt_acc.f90:756-759), k-range
 * ceiling(Kmax/3.) (:760-762), Standard_Ewald + real_space (:768-770);
 * against SPME + real_space at the production parameters (:783-793);
 * rmsf = sqrt(sum |f_se - f_c|^2 / sum |f_se|^2) over every particle
 * (the reference sums rows 1..Nq, :797-801; corrected).
 * out: rmsf, f_r, f_k, alpha', rcc', K'1, K'2, K'3.  acc_se, acc_c (NN,3)
 * may be null. */
void orc_error_analysis(orc_sys *s, double out[8], double *acc_se, double *acc_c) {
  md_params *p = &s->p;
  const int NN = s->NN;
  double qtot = 0;
  for (int m = 0; m < s->Nq; m++) qtot += QC(s->charge[m]) * QC(s->charge[m]);
  const double third = (double)(1.0f / 3.0f), V = p->Lx * p->Ly * p->Lz;
  const double ex = exp(-p->tol * p->tol);
  const double f_r = 2 * fabs(p->qq) * sqrt(qtot / p->rcc / p->Lx / p->Ly / p->Lz) * ex;
  const float kprod = powf((float)(p->ng[0] * p->ng[1] * p->ng[2]), 1.0f / 3.0f);
  const double f_k = fabs(p->qq) * p->alpha / pow(V, third) / PI * sqrt(8 * qtot / (double)kprod) * ex;
  md_params prod = *p, ref = *p;
  ref.ovr_alpha = ref.ovr_rcc = 0;
  ref.ovr_ng[0] = ref.ovr_ng[1] = ref.ovr_ng[2] = 0;
  orc_ewald_rule(&ref, 5.0, 10.0);
  int K[3];
  for (int d = 0; d < 3; d++) K[d] = (int)ceilf((float)ref.ng[d] / 3.0f);
  double *fse = calloc((size_t)NN * 3, sizeof(double));
  double *fc = calloc((size_t)NN * 3, sizeof(double));
  /* every pair within rcc', the set the reference's list or pair loop
   * visits, in the pair-loop form */
  ref.real_verlet = 0;
  s->p = ref;
  orc_standard_ewald(s, ref.alpha, K[0], K[1], K[2], fse);
  orc_real_space(s, fse);
  s->p = prod;
  orc_spme(s, fc);
  orc_real_space(s, fc);
  double s1 = 0, s2 = 0;
  for (size_t k = 0; k < (size_t)NN * 3; k++) {
    s1 += (fse[k] - fc[k]) * (fse[k] - fc[k]);
    s2 += fse[k] * fse[k];
  }
  out[0] = s2 > 0 ? sqrt(s1 / s2) : 0;
  out[1] = f_r;
  out[2] = f_k;
  out[3] = ref.alpha;
  out[4] = ref.rcc;
  out[5] = K[0];
  out[6] = K[1];
  out[7] = K[2];
  if (acc_se) memcpy(acc_se, fse, sizeof(double) * NN * 3);
  if (acc_c) memcpy(acc_c, fc, sizeof(double) * NN * 3);
  free(fse);
  free(fc);
}

/* ------------------------------------------------------------ energies */

/* The reference computes only the kinetic energy (height,
 * in_output.f90:1307-1311); the potentials below are the ones whose negative
 * gradients are the reference's forces, with the same lists, cuts, classes
 * and constants as orc_lj_force, orc_fene_force, orc_real_space and
 * orc_spme:
 *   out[0]  WCA: 4 ((s/r)^12 - (s/r)^6) + 1 per pair within the class cut
 *           (compt_acc.f90:922-952; the +1 puts the energy at 0 at
 *           r = 2^(1/6) s), half per entry of the full list
 *   out[1]  wall: 0.4/3 ((s/d)^9 - (s/d)^3) on the wall the force acts from
 *           (compt_acc.f90:955-975)
 *   out[2]  FENE: -k R0^2 / 2 ln(1 - r^2/R0^2) per bond (compt_acc.f90:865-892)
 *   out[3]  real space: lB/Beta q_i q_j erfc(a r)/r per pair within r_c
 *           (compt_acc.f90:997-1038; a full-list entry counts half)
 *   out[4]  reciprocal: lB/Beta / (2 K1 K2 K3) sum_mesh Q U, with U the
 *           unnormalised inverse FFT of BC times the FFT of Q
 *           (compt_acc.f90:1368-1398, whose gradient is IFrc :1402-1467)
 *   out[5]  slab: 2 pi lB/Beta (sum q z)^2 / (Lx Ly Lz Zfac)
 *           (the potential of the correction at compt_acc.f90:1040-1048)
 *   out[6]  Ewald self term: -lB/Beta a/sqrt(pi) sum q^2
 *   out[7]  out[3] + out[4] + out[5] + out[6]
 *   out[8]  field: -E_z sum q z (compt_acc.f90:855)
 *   out[9]  potential total: out[0] + out[1] + out[2] + out[7] + out[8]
 *   out[10] kinetic: sum |v|^2 / 2
 *   out[11] the reference's kenetic_energy: out[10] / (NN - N_anchor)
 *   out[12] LJ pair virial sum r.F, out[13] FENE, out[14] real space
 *   out[15] 0
 * The reference's ZeroForce (mean reciprocal force removed) has no energy
 * counterpart. */
void orc_energies(orc_sys *s, double *out) {
  md_params *p = &s->p;
  int NN = s->NN;
  double Lx = p->Lx, Ly = p->Ly, Lz = p->Lz;
  double hLx = Lx / 2, nhLx = -Lx / 2, hLy = Ly / 2, nhLy = -Ly / 2;
  double rcl2 = p->rcl * p->rcl, rcl3 = p->rcl_s * p->rcl_s, rcl4 = p->rcl_sp * p->rcl_sp;
  double sigma2 = p->sigma_s * p->sigma_s, sigma3 = p->sigma_sp * p->sigma_sp;
  memset(out, 0, sizeof(double) * 16);
  for (int64_t k = 0; k < s->npair1; k++) {
    int i = s->l1i[k], j = s->l1j[k];
    double rij[3] = {X(i) - X(j), Y(i) - Y(j), Z(i) - Z(j)};
    int m = (int)(fabs(QC(i)) + fabs(QC(j)));
    if (rij[0] > hLx) rij[0] = rij[0] - Lx;
    else if (rij[0] <= nhLx) rij[0] = rij[0] + Lx;
    if (rij[1] > hLy) rij[1] = rij[1] - Ly;
    else if (rij[1] <= nhLy) rij[1] = rij[1] + Ly;
    double rsqr = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    double inv_r2 = 0;
    int hit = 0;
    if (m == 6) {
      if (rsqr < rcl3) { inv_r2 = sigma2 / rsqr; hit = 1; }
    } else if (m < 3) {
      if (rsqr < rcl2) { inv_r2 = 1 / rsqr; hit = 1; }
    } else {
      if (rsqr < rcl4) { inv_r2 = sigma3 / rsqr; hit = 1; }
    }
    if (hit) {
      double inv_r6 = inv_r2 * inv_r2 * inv_r2;
      out[0] += 0.5 * (4 * (inv_r6 * inv_r6 - inv_r6) + 1);
      out[12] += 0.5 * 48 * inv_r6 * (inv_r6 - 0.5);
    }
  }
  const double c04 = (double)0.4f;
  const double c5 = (double)(float)pow(5.0, (double)(1.0f / 6.0f));
  for (int i = 0; i < NN; i++) {
    double z = Z(i);
    if (fabs(z) < 0.01 || fabs(z - Lz) < 0.01) continue;
    double ss = QC(i) != 3 ? 1.0 : p->sigma_s;
    double rw = (QC(i) != 3 ? p->rcl : p->rcl_s) / c5, dl = -1;
    if (z < rw) dl = z;
    else if ((Lz - z) < rw) dl = Lz - z;
    if (dl > 0) out[1] += c04 / 3 * (pow(ss / dl, 9) - pow(ss / dl, 3));
  }
  for (int b = 0; b < s->N_bond; b++) {
    int j = s->fene[b], k = s->fene[s->N_bond + b];
    double r[3];
    double rsqr = rij_rr(s, j, k, r);
    out[2] += -0.5 * p->kFENE * p->R0_2 * log(1 - rsqr / p->R0_2);
    out[13] += -p->kFENE * p->R0_2 / (p->R0_2 - rsqr) * rsqr;
  }
  if (p->qq != 0 && s->Nq > 0) {
    double c1 = p->lb / p->Beta, c2 = 2 * p->alpha / sqrt(PI);
    double alpha2 = p->alpha * p->alpha, rcc2 = p->rcc * p->rcc;
    if (p->real_verlet == 0) {
      for (int m = 0; m < s->Nq - 1; m++)
        for (int n = m + 1; n < s->Nq; n++) {
          int i = s->charge[m], j = s->charge[n];
          double rij[3];
          double rsqr = rij_rr(s, i, j, rij);
          if (rsqr < rcc2) {
            double rr = sqrt(rsqr);
            out[3] += c1 * QC(i) * QC(j) * erfc(p->alpha * rr) / rr;
            out[14] += c1 * QC(i) * QC(j) * (erfc(p->alpha * rr) / rr + c2 * exp(-alpha2 * rsqr));
          }
        }
    } else {
      for (int64_t k = 0; k < s->npair2; k++) {
        int i = s->l2i[k], j = s->l2j[k];
        double rij[3];
        double rsqr = rij_rr(s, i, j, rij);
        if (rsqr < rcc2) {
          double rr = sqrt(rsqr);
          out[3] += 0.5 * c1 * QC(i) * QC(j) * erfc(p->alpha * rr) / rr;
          out[14] += 0.5 * c1 * QC(i) * QC(j) * (erfc(p->alpha * rr) / rr + c2 * exp(-alpha2 * rsqr));
        }
      }
    }
    double *tmp = calloc((size_t)NN * 3, sizeof(double));
    orc_spme(s, tmp);
    free(tmp);
    size_t M = (size_t)p->ng[0] * p->ng[1] * p->ng[2];
    double qu = 0;
    for (size_t g = 0; g < M; g++) qu += creal(s->Q[g]) * creal(s->U[g]);
    out[4] = c1 * qu / (2.0 * (double)M);
    double qz = 0, q2 = 0;
    for (int m = 0; m < s->Nq; m++) {
      int i = s->charge[m];
      qz += QC(i) * Z(i);
      q2 += QC(i) * QC(i);
    }
    out[5] = 2 * PI * c1 * qz * qz / (Lx * Ly * Lz * p->Zfac);
    out[6] = -c1 * p->alpha / sqrt(PI) * q2;
  }
  out[7] = out[3] + out[4] + out[5] + out[6];
  for (int i = 0; i < NN; i++) out[8] += -p->EF * QC(i) * Z(i);
  out[9] = out[0] + out[1] + out[2] + out[7] + out[8];
  if (s->vel) {
    for (int i = 0; i < NN; i++)
      for (int d = 0; d < 3; d++) out[10] += s->vel[(size_t)d * NN + i] * s->vel[(size_t)d * NN + i] / 2.;
  }
  out[11] = NN > s->N_anchor ? out[10] / (NN - s->N_anchor) : 0;
}

/* ------------------------------------------------------------- noise */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Three standard normals for particle i (0-based) at `step`: one
 * Philox4x32-10 block, counter (i, step, 0x4D444E53, 0), key = seed; words
 * mapped to (0,1) as (w+0.5)/2^32; Box-Muller, both branches of the first
 * pair, cosine branch of the second (the reference draws its noise with
 * gauss_dist, glob_var.f90:108-121, from an OS-seeded stream). */
void orc_noise_normals(uint64_t seed, int32_t step, int32_t i, double n[3]) {
  uint32_t ctr[4] = {(uint32_t)i, (uint32_t)step, 0x4D444E53u, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  const double s32 = 1.0 / 4294967296.0;
  double u0 = ((double)w[0] + 0.5) * s32, u1 = ((double)w[1] + 0.5) * s32;
  double u2 = ((double)w[2] + 0.5) * s32, u3 = ((double)w[3] + 0.5) * s32;
  double r1 = sqrt(-2 * log(u0)), r2 = sqrt(-2 * log(u2));
  n[0] = r1 * cos(2 * PI * u1);
  n[1] = r1 * sin(2 * PI * u1);
  n[2] = r2 * cos(2 * PI * u3);
}

/* ------------------------------------------------------------ stepping */

/* Compute_Force (compt_acc.f90:815-862). */
int orc_compute_force(orc_sys *s, int32_t step, int32_t noise,
                      const double *eta) {
  md_params *p = &s->p;
  int NN = s->NN;
  double *tmp = malloc(sizeof(double) * 3 * NN);
  if (p->qq != 0 && step % p->multistep == 0) {
    memset(s->acc_c, 0, sizeof(double) * 3 * NN);
    orc_spme(s, s->acc_c);
    orc_real_space(s, s->acc_c);
  }
  memcpy(s->acc, s->acc_c, sizeof(double) * 3 * NN);
  orc_lj_force(s, tmp);
#pragma omp parallel for
  for (int i = 0; i < 3 * NN; i++) s->acc[i] = s->acc[i] + tmp[i];
  int rc = orc_fene_force(s, tmp);
  if (rc != MD_OK) {
    free(tmp);
    return rc;
  }
#pragma omp parallel for
  for (int i = 0; i < 3 * NN; i++) s->acc[i] = s->acc[i] + tmp[i];
  double sig = sqrt(2.0 / p->Beta / p->dt * p->xi);
#pragma omp parallel for
  for (int i = 0; i < NN; i++) {
    double n[3] = {0, 0, 0};
    if (noise == MD_NOISE_ETA) {
      n[0] = eta[i]; n[1] = eta[NN + i]; n[2] = eta[2 * (size_t)NN + i];
    } else if (noise == MD_NOISE_PHILOX) {
      orc_noise_normals(p->seed, step, i, n);
    }
    s->acc[i] = s->acc[i] + n[0] * sig;
    s->acc[NN + i] = s->acc[NN + i] + n[1] * sig;
    s->acc[2 * (size_t)NN + i] = s->acc[2 * (size_t)NN + i] + n[2] * sig + QC(i) * p->EF;
  }
  for (int a = 0; a < s->N_anchor; a++) {
    int j = s->anchor[a];
    s->acc[j] = s->acc[NN + j] = s->acc[2 * (size_t)NN + j] = 0;
  }
  free(tmp);
  return MD_OK;
}

/* new_position (init_update.f90:555-607) with period_condition_pos
 * (init_update.f90:610-643). */
int orc_new_position(orc_sys *s, int32_t step, int32_t noise,
                     const double *eta) {
  md_params *p = &s->p;
  int NN = s->NN;
  double dt = p->dt, xi = p->xi;
#pragma omp parallel for
  for (int i = 0; i < 3 * NN; i++)
    s->vel[i] = s->acc[i] * dt / 2 + (1 - xi * dt / 2) * s->vel[i];
  double dmax1 = 0, dmax2 = 0;
#pragma omp parallel for reduction(max : dmax1, dmax2)
  for (int i = 0; i < NN; i++) {
    double dri[3];
    for (int d = 0; d < 3; d++) {
      dri[d] = s->vel[(size_t)d * NN + i] * dt;
      s->pos[(size_t)d * NN + i] = s->pos[(size_t)d * NN + i] + dri[d];
    }
    double dr1 = sqrt(dri[0] * dri[0] + dri[1] * dri[1] + dri[2] * dri[2]);
    if (dmax1 <= dr1) dmax1 = dr1;
    if (QC(i) != 0)
      if (dmax2 <= dr1) dmax2 = dr1;
  }
  s->dr_max1 = s->dr_max1 + dmax1;
  s->dr_max2 = s->dr_max2 + dmax2;
  for (int k = 0; k < NN; k++) {
    if (X(k) > p->Lx / 2) X(k) = X(k) - p->Lx;
    else if (X(k) <= -p->Lx / 2) X(k) = X(k) + p->Lx;
    if (Y(k) > p->Ly / 2) Y(k) = Y(k) - p->Ly;
    else if (Y(k) <= -p->Ly / 2) Y(k) = Y(k) + p->Ly;
  }
  int rc = orc_compute_force(s, step, noise, eta);
#pragma omp parallel for
  for (int i = 0; i < 3 * NN; i++)
    s->vel[i] = (s->vel[i] + s->acc[i] * dt / 2.) / (1 + xi * dt / 2.);
  return rc;
}

/* update_verlet_list (compt_acc.f90:699-726). */
int orc_update_verlet_list(orc_sys *s) {
  if (s->dr_max1 > (s->p.rvl_eff - s->p.rcl_s) / 2) {
    if (orc_lj_verlet_list(s) < 0) return MD_ECELL;
    s->dr_max1 = 0;
  }
  if (s->dr_max2 > s->p.rsk / 2 && s->p.real_verlet != 0) {
    if (orc_real_verlet_list(s) < 0) return MD_ECELL;
    s->dr_max2 = 0;
  }
  return MD_OK;
}

/* rescale_velocity (init_update.f90:498-552). */
void orc_rescale_velocity(orc_sys *s) {
  int NN = s->NN;
  double v_sum[3] = {0, 0, 0}, v2_sum = 0;
  for (int i = 0; i < NN; i++)
    for (int d = 0; d < 3; d++) v_sum[d] = v_sum[d] + s->vel[(size_t)d * NN + i];
  for (int d = 0; d < 3; d++) v_sum[d] = v_sum[d] / (NN - s->N_anchor);
  for (int i = 0; i < NN; i++) {
    for (int d = 0; d < 3; d++) s->vel[(size_t)d * NN + i] -= v_sum[d];
    double vx = s->vel[i], vy = s->vel[NN + i], vz = s->vel[2 * (size_t)NN + i];
    v2_sum = v2_sum + vx * vx + vy * vy + vz * vz;
  }
  for (int a = 0; a < s->N_anchor; a++) {
    int j = s->anchor[a];
    double vx = s->vel[j], vy = s->vel[NN + j], vz = s->vel[2 * (size_t)NN + j];
    v2_sum = v2_sum - (vx * vx + vy * vy + vz * vz);
    s->vel[j] = s->vel[NN + j] = s->vel[2 * (size_t)NN + j] = 0;
  }
  v2_sum = v2_sum / (NN - s->N_anchor - 1);
  double fac = sqrt(3.0 / s->p.Beta / v2_sum);
  for (int i = 0; i < 3 * NN; i++) s->vel[i] = s->vel[i] * fac;
}

/* The step loop of main.f90:67-94 with Philox noise. */
int orc_run(orc_sys *s, int32_t step0, int32_t nsteps, int32_t preheat) {
  for (int32_t step = step0; step < step0 + nsteps; step++) {
    int rc = orc_new_position(s, step, MD_NOISE_PHILOX, NULL);
    if (rc) return rc;
    if (preheat && step % 20 == 0 && step < 300000) orc_rescale_velocity(s);
    rc = orc_update_verlet_list(s);
    if (rc) return rc;
  }
  return MD_OK;
}
