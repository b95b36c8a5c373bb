// md_host.cpp -- host-side set-up of the brush engine: parameter files,
// derived quantities, topology, seeded initial configuration.  Runs once per
// system on the CPU; nothing here is on the step path.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/md_engine.h"
#include "md_kernels.h"

namespace {

constexpr double kPi = 3.141592653589793;

int round_half_away(double x) { return (int)std::lround(x); }

// First token of each non-empty line, Fortran D exponents accepted.
bool read_values(const char *path, std::vector<double> &out, size_t need) {
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  char line[512];
  while (out.size() < need && std::fgets(line, sizeof line, f)) {
    char *s = line;
    while (*s && std::isspace((unsigned char)*s)) s++;
    if (!*s) continue;
    char tok[128];
    size_t n = 0;
    while (s[n] && !std::isspace((unsigned char)s[n]) && s[n] != '!' && s[n] != ',' &&
           n < sizeof tok - 1) {
      tok[n] = (s[n] == 'D' || s[n] == 'd') ? 'E' : s[n];
      n++;
    }
    tok[n] = 0;
    if (n == 0) continue;
    char *end = nullptr;
    double v = std::strtod(tok, &end);
    if (end == tok) {
      std::fclose(f);
      return false;
    }
    out.push_back(v);
  }
  std::fclose(f);
  return out.size() == need;
}

// Sequential uniform stream for set-up: Philox4x32-10 blocks with counter
// (block, 0, tag, 0); 53-bit uniforms in (0,1).
struct SetupRng {
  uint64_t seed;
  uint32_t tag;
  uint32_t block = 0;
  uint32_t w[4];
  int used = 4;
  double next() {
    if (used >= 4) {
      mdk::philox4x32_10(block++, 0u, tag, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
      used = 0;
    }
    uint32_t a = w[used], b = w[used + 1];
    used += 2;
    return ((double)a * 2097152.0 + (double)(b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace

extern "C" int md_read_params(const char *system_data, const char *force_data,
                              const char *q_arm, md_params *p) {
  if (!p || !system_data || !force_data) return MD_EINVAL;
  std::memset(p, 0, sizeof *p);
  std::vector<double> s, f;
  if (!read_values(system_data, s, 23)) return MD_EIO;
  if (!read_values(force_data, f, 16)) return MD_EIO;
  // system_data.txt order: in_output.f90:190-212
  p->uniform_or_random = (int)s[0];
  p->Lz = s[1];
  p->Z_empty = s[2];
  p->sigmag = s[3];
  p->Beta = s[4];
  p->qq = s[5];
  p->qqi = s[6];
  p->ion_ratio = s[7];
  p->arm = (int)s[8];
  p->Nma = (int)s[9];
  p->Nga = (int)s[10];
  p->Nml = (int)s[11];
  p->Ngl = (int)s[12];
  p->man_l = (int)s[13];
  p->man_s = (int)s[14];
  p->R_bond = s[15];
  p->StepNum0 = (int)s[16];
  p->StepNum = (int)s[17];
  p->DeltaStep1 = (int)s[18];
  p->DeltaStep2 = (int)s[19];
  p->DeltaStep3 = (int)s[20];
  p->multistep = (int)s[21];
  p->dt = s[22];
  // force_data.txt order: compt_acc.f90:141-156
  p->epsilon = f[0];
  p->sigma = f[1];
  p->sigma_s = f[2];
  p->rcl = f[3];
  p->rvl = f[4];
  p->rsk = f[5];
  p->R0_2 = f[6];
  p->kFENE = f[7];
  p->ordr[0] = (int)f[8];
  p->ordr[1] = (int)f[9];
  p->ordr[2] = (int)f[10];
  p->lb = f[11];
  p->xi = f[12];
  p->EF = f[13];
  p->tol = f[14];
  p->tau_rf = f[15];
  if (p->arm < 0 || p->arm > MD_MAX_ARM) return MD_EINVAL;
  if (p->Nga != 0) {
    std::vector<double> a;
    if (!q_arm || !read_values(q_arm, a, (size_t)p->arm)) return MD_EIO;
    for (int i = 0; i < p->arm; i++) p->charged_arm[i] = (int)a[i];
  }
  return MD_OK;
}

namespace mdk {

// Initialize_ewald_parameters (compt_acc.f90:387-414) at (tol, tau_rf):
// alpha, r_c, r_v, the k-ranges before the mesh factor 3, and whether the
// real space keeps a Verlet list; the overrides ovr_alpha / ovr_rcc apply
// when use_ovr.
void ewald_rule(const md_params *p, double tol, double tau_rf, bool use_ovr, double *alpha_out,
                double *rcc_out, double *rvc_out, int k[3], int *real_verlet) {
  const double V = p->Lx * p->Ly * p->Lz;
  double alpha = std::pow(tau_rf * kPi * kPi * kPi * p->Nq / (V * V), 1.0 / 6);
  double rcc = tol / alpha;
  if (use_ovr && p->ovr_alpha > 0) {
    alpha = p->ovr_alpha;
    rcc = p->ovr_rcc > 0 ? p->ovr_rcc : tol / alpha;
  }
  double rvc = rcc + p->rsk;
  const double len[3] = {p->Lx, p->Ly, p->Lz * p->Zfac};
  if ((int)(p->Lx / rvc) * (int)(p->Ly / rvc) * (int)(p->Lz / rvc) > 27) {
    for (int d = 0; d < 3; d++) k[d] = (int)std::ceil(tol * len[d] * alpha / kPi);
    *real_verlet = 1;
  } else {
    rcc = std::min(p->Lx, p->Ly) / 2;
    if (p->Lx <= p->Ly) rcc = p->Lx / 2;
    for (int d = 0; d < 3; d++) k[d] = (int)std::ceil(tol * tol / kPi * len[d] / rcc);
    alpha = tol / rcc;
    rvc = rcc + p->rsk;
    *real_verlet = 0;
  }
  *alpha_out = alpha;
  *rcc_out = rcc;
  *rvc_out = rvc;
}

}  // namespace mdk

extern "C" int md_derive_params(md_params *p) {
  if (!p || p->arm < 0 || p->arm > MD_MAX_ARM || p->sigmag <= 0 || p->Lz <= 0)
    return MD_EINVAL;
  // particle counts: in_output.f90:250-283
  int charged_arms = 0;
  if (p->Nga != 0)
    for (int a = 0; a < p->arm; a++) charged_arms += p->charged_arm[a];
  p->Nta = (p->arm * p->Nma + 1) * p->Nga;
  p->Ntl = p->Ngl * p->Nml;
  p->Npe = p->Nta + p->Ntl;
  p->N_anchor = p->Nga + p->Ngl;
  if (p->N_anchor <= 0) return MD_EINVAL;
  int zq = round_half_away(std::fabs(p->qq)), zi = round_half_away(std::fabs(p->qqi));
  p->Nq_PE = 0;
  if (p->qq != 0 && p->man_l != 0 && p->man_s != 0) {
    // floor of a default-real quotient of two integers
    int lin = (int)std::floor((float)(p->Nml - 1) / (float)p->man_l);
    int star = (int)std::floor((float)p->Nma / (float)p->man_s);
    p->Nq_PE = lin * p->Ngl + star * charged_arms * p->Nga;
  }
  p->Nq_salt_ions = 0;
  if (p->qqi != 0) p->Nq_salt_ions = round_half_away(p->ion_ratio * p->Nq_PE * zq) / zi;
  p->Nq = p->Nq_PE * (zq + 1) + p->Nq_salt_ions * (zi + 1);
  p->NN = p->Npe + p->Nq_PE * zq + p->Nq_salt_ions * (zi + 1);
  p->N_bond = p->arm * p->Nma * p->Nga + (p->Nml - 1) * p->Ngl;
  // box: in_output.f90:286-288
  p->Lx = std::sqrt(p->N_anchor / p->sigmag);
  p->Ly = p->Lx;
  p->Zfac = (p->Lz + p->Lx * p->Z_empty) / p->Lz;
  // radii: compt_acc.f90:159-162, with 2**(1./6) in default real
  const float two6 = (float)std::pow(2.0, (double)(1.0f / 6.0f));
  p->sigma_sp = (p->sigma + p->sigma_s) / 2;
  p->rcl_s = (double)two6 * p->sigma_s;
  p->rcl_sp = (double)two6 * p->sigma_sp;
  // LJ list radius rcl_s + 1 (compt_acc.f90:162), or the skin a caller chose
  p->rvl_eff = p->rcl_s + (p->ovr_skin > 0 ? 2 * p->ovr_skin : 1);
  p->gdim[0] = p->Lx;
  p->gdim[1] = p->Ly;
  p->gdim[2] = p->Lz * p->Zfac;
  p->real_verlet = 0;
  p->alpha = p->rcc = p->rvc = 0;
  p->ng[0] = p->ng[1] = p->ng[2] = 0;
  if (p->qq == 0) return MD_OK;
  double alpha, rcc, rvc;
  int k[3];
  mdk::ewald_rule(p, p->tol, p->tau_rf, true, &alpha, &rcc, &rvc, k, &p->real_verlet);
  p->alpha = alpha;
  p->rcc = rcc;
  p->rvc = rvc;
  for (int d = 0; d < 3; d++) {
    int kk = (int)std::ceil((float)k[d] * 3.0f);
    p->ng[d] = p->ovr_ng[d] > 0 ? p->ovr_ng[d] : kk;
  }
  return MD_OK;
}

extern "C" int md_build_topology(const md_params *p, int32_t *charge, int32_t *fene,
                                 int32_t *anchor) {
  if (!p || !fene || !anchor) return MD_EINVAL;
  const int per_star = p->arm * p->Nma + 1;
  const int nb = p->N_bond;
  // anchors: compt_acc.f90:337-342
  for (int s = 0; s < p->Nga; s++) anchor[s] = s * per_star + 1;
  for (int c = 0; c < p->Ngl; c++) anchor[p->Nga + c] = p->Nta + c * p->Nml + 1;
  // bonds: compt_acc.f90:345-371
  int b = 0;
  auto bond = [&](int i, int j) {
    fene[b] = i;
    fene[nb + b] = j;
    b++;
  };
  for (int s = 0; s < p->Nga; s++) {
    const int base = s * per_star;
    for (int k = 1; k <= p->Nma; k++) bond(base + k, base + k + 1);  // anchor .. branch
    for (int a = 2; a <= p->arm; a++) {
      const int first = base + (a - 1) * p->Nma + 1;  // bead before arm a
      bond(base + p->Nma + 1, first + 1);
      for (int k = 1; k < p->Nma; k++) bond(first + k, first + k + 1);
    }
  }
  for (int c = 0; c < p->Ngl; c++)
    for (int k = 1; k < p->Nml; k++) bond(p->Nta + c * p->Nml + k, p->Nta + c * p->Nml + k + 1);
  if (b != nb) return MD_EINVAL;
  if (!charge || p->qq == 0) return MD_OK;
  // charged beads then ions: compt_acc.f90:452-487
  int m = 0;
  for (int s = 0; s < p->Nga; s++) {
    const int base = s * per_star;
    if (p->charged_arm[0] != 0)
      for (int k = 2; k <= p->Nma + 1; k++)
        if ((k - 1) % p->man_s == 0) charge[m++] = base + k;
    for (int a = 2; a <= p->arm; a++) {
      if (p->charged_arm[a - 1] == 0) continue;
      for (int k = 1; k <= p->Nma; k++)
        if (k % p->man_s == 0) charge[m++] = base + (a - 1) * p->Nma + 1 + k;
    }
  }
  for (int c = 0; c < p->Ngl; c++)
    for (int k = 2; k <= p->Nml; k++)
      if ((k - 1) % p->man_l == 0) charge[m++] = p->Nta + c * p->Nml + k;
  for (int i = p->Npe + 1; i <= p->NN; i++) charge[m++] = i;
  return m == p->Nq ? MD_OK : MD_EINVAL;
}

// ---------------------------------------------------------------------------
// Initial configuration: init_update.f90:6-458 with a seeded Philox stream.
// Beads are placed in topology order (stars 1..Nta, then linear chains);
// the reference's uniform mode writes linear chains from index 1 and then
// overwrites them with stars (init_update.f90:38-39,263), which only works
// when one of the two kinds is absent, so the index ranges are fixed here.
namespace {

struct Builder {
  const md_params *p;
  double *x, *y, *z, *q;
  int placed = 0;
  SetupRng rng;
  int status = MD_OK;

  void wrap(double &a, double L) {  // period_condition_rij, init_update.f90:646-677
    if (a > L / 2) a -= L;
    else if (a <= -L / 2) a += L;
  }
  double dist2(int n, double px, double py, double pz) {  // rij_and_rr
    double dx = x[n] - px, dy = y[n] - py, dz = z[n] - pz;
    if (dx > p->Lx / 2) dx -= p->Lx;
    else if (dx <= -p->Lx / 2) dx += p->Lx;
    if (dy > p->Ly / 2) dy -= p->Ly;
    else if (dy <= -p->Ly / 2) dy += p->Ly;
    return dx * dx + dy * dy + dz * dz;
  }
  // rejection: init_update.f90:143-150 and :382-389
  bool clash(int l, double px, double py, double pz) {
    if (pz < 0.9) return true;
    for (int n = 0; n < l; n++)
      if (dist2(n, px, py, pz) < 0.7) return true;
    return false;
  }
  // one bead bonded to `from`; `tilt` bends the first ten tries upward
  void grow(int l, int from, int mode) {
    const double R = p->R_bond;
    for (int tries = 0;; tries++) {
      if (tries > 200000) {
        status = MD_EINVAL;
        return;
      }
      double r1 = rng.next(), r2 = rng.next();
      if (mode == 1) {  // linear chain: init_update.f90:368-372
        r1 = tries < 10 ? r1 / 5 : r1 * r1;
      } else if (mode == 2) {  // star arm: init_update.f90:294-296
        if (tries < 10) r1 = r1 / 1000;
      }  // mode 0: branch bead, full sphere (init_update.f90:289-292)
      double px = x[from] + R * std::sin(kPi * r1) * std::cos(2 * kPi * r2);
      double py = y[from] + R * std::sin(kPi * r1) * std::sin(2 * kPi * r2);
      double pz = z[from] + R * std::cos(kPi * r1);
      wrap(px, p->Lx);
      wrap(py, p->Ly);
      if (!clash(l, px, py, pz)) {
        x[l] = px; y[l] = py; z[l] = pz;
        return;
      }
    }
  }
  bool anchor_ok(int l, double px, double py) {  // init_update.f90:105-113
    for (int n = 0; n < l; n++)
      if (dist2(n, px, py, 0.0) < 0.8) return false;
    return true;
  }
};

}  // namespace

extern "C" int md_initialize_position(const md_params *p, uint64_t seed, double *pos) {
  if (!p || !pos || p->NN <= 0) return MD_EINVAL;
  const int NN = p->NN;
  Builder B{p, pos, pos + NN, pos + 2 * (size_t)NN, pos + 3 * (size_t)NN};
  B.rng.seed = seed;
  B.rng.tag = mdk::kInitTag;
  std::memset(pos, 0, sizeof(double) * 4 * (size_t)NN);
  const int na = p->N_anchor;
  const int side = (int)std::lround(std::sqrt((double)na));
  // grafting sites
  std::vector<double> sx, sy;
  std::vector<int> kind;  // 0 star, 1 linear
  if (p->uniform_or_random == 0) {
    // square lattice, checkerboard when the two kinds are equal in number
    // (init_update.f90:267-279, 347-359)
    for (int i = 1; i <= na; i++) {
      int ix = (i - 1) / side + 1, iy = (i - 1) % side + 1;
      sx.push_back(p->Lx / side * (ix - 0.5) - p->Lx / 2);
      sy.push_back(p->Ly / side * (iy - 0.5) - p->Ly / 2);
      kind.push_back(p->Nga == p->Ngl ? ((ix + iy) % 2 == 0 ? 1 : 0) : (p->Nga > 0 ? 0 : 1));
    }
  }
  int used_star = 0, used_lin = 0;
  auto site = [&](int want, double &ax, double &ay) {
    if (p->uniform_or_random == 0) {
      int &cursor = want == 0 ? used_star : used_lin;
      while (cursor < na && kind[cursor] != want) cursor++;
      if (cursor < na) {
        ax = sx[cursor];
        ay = sy[cursor];
        cursor++;
        return;
      }
      // lattice ran out (uneven split): fall back to random grafting
    }
    for (int tries = 0; tries < 200000; tries++) {  // init_update.f90:97-114
      double cx = B.rng.next() * p->Lx - p->Lx / 2, cy = B.rng.next() * p->Ly - p->Ly / 2;
      if (B.anchor_ok(B.placed, cx, cy)) {
        ax = cx;
        ay = cy;
        return;
      }
    }
    B.status = MD_EINVAL;
  };
  int l = 0;
  for (int s = 0; s < p->Nga && B.status == MD_OK; s++) {  // stars
    site(0, B.x[l], B.y[l]);
    B.z[l] = 0;
    const int anchor = l++;
    B.placed = l;
    for (int a = 1; a <= p->arm && B.status == MD_OK; a++)
      for (int k = 1; k <= p->Nma && B.status == MD_OK; k++) {
        int from = (k == 1 && a != 1) ? anchor + p->Nma : l - 1;
        B.grow(l, from, (k == 1 && a != 1) ? 0 : 2);
        B.placed = ++l;
      }
  }
  for (int c = 0; c < p->Ngl && B.status == MD_OK; c++) {  // linear chains
    site(1, B.x[l], B.y[l]);
    B.z[l] = 0;
    B.placed = ++l;
    for (int k = 2; k <= p->Nml && B.status == MD_OK; k++) {
      B.grow(l, l - 1, 1);
      B.placed = ++l;
    }
  }
  if (B.status != MD_OK) return B.status;
  // ions: init_update.f90:419-456
  const int salt = p->Nq_salt_ions;
  const int zi = (int)std::lround(std::fabs(p->qqi));
  for (int i = p->Npe; i < NN; i++) {
    const bool is_salt_i = (i + 1) > NN - salt;
    bool ok = false;
    for (int tries = 0; tries < 1000000 && !ok; tries++) {
      double px = B.rng.next() * p->Lx - p->Lx / 2;
      double py = B.rng.next() * p->Ly - p->Ly / 2;
      double pz = B.rng.next() * (p->Lz - 1.8) + 0.9;
      ok = true;
      for (int j = 0; j < i && ok; j++) {
        const bool is_salt_j = (j + 1) > NN - salt;
        double r2 = B.dist2(j, px, py, pz);
        if (!is_salt_i && !is_salt_j) ok = r2 >= 0.8;
        else if (is_salt_i && !is_salt_j) ok = r2 >= p->sigma_s * 0.8;
        else if (is_salt_i && is_salt_j) ok = r2 >= p->sigma_s * p->sigma_s * 0.8;
      }
      if (ok) {
        B.x[i] = px; B.y[i] = py; B.z[i] = pz;
      }
    }
    if (!ok) return MD_EINVAL;
    const int n1 = i + 1;
    if (n1 <= NN - salt * (zi + 1)) B.q[i] = -p->qq / std::fabs(p->qq);
    else if (n1 <= NN - salt) B.q[i] = -p->qqi / std::fabs(p->qqi);
    else B.q[i] = p->qqi;
  }
  // charges on the chains: init_update.f90:52-54
  if (p->qq != 0 && p->Nq > 0) {
    std::vector<int32_t> charge(p->Nq), fene(2 * (size_t)std::max(p->N_bond, 1)),
        anchor(std::max(p->N_anchor, 1));
    int rc = md_build_topology(p, charge.data(), fene.data(), anchor.data());
    if (rc) return rc;
    for (int m = 0; m < p->Nq_PE; m++) B.q[charge[m] - 1] = p->qq;
  }
  return MD_OK;
}

extern "C" int md_initialize_velocity(const md_params *p, uint64_t seed, double *vel) {
  if (!p || !vel) return MD_EINVAL;
  const int NN = p->NN;
  SetupRng rng{seed, mdk::kVelTag};
  // gauss_dist(0, sqrt(1/Beta)) per component: init_update.f90:482-487
  const double sd = std::sqrt(1.0 / p->Beta);
  for (int i = 0; i < NN; i++)
    for (int d = 0; d < 3; d++) {
      double u1 = rng.next(), u2 = rng.next();
      vel[(size_t)d * NN + i] = std::sqrt(-2 * std::log(u1)) * std::cos(2 * kPi * u2) * sd;
    }
  std::vector<int32_t> fene(2 * (size_t)std::max(p->N_bond, 1)), anchor(std::max(p->N_anchor, 1));
  int rc = md_build_topology(p, nullptr, fene.data(), anchor.data());
  if (rc) return rc;
  for (int a = 0; a < p->N_anchor; a++)
    for (int d = 0; d < 3; d++) vel[(size_t)d * NN + anchor[a] - 1] = 0;
  // rescale_velocity: init_update.f90:522-551
  double vs[3] = {0, 0, 0}, v2 = 0;
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < NN; i++) vs[d] += vel[(size_t)d * NN + i];
  for (int d = 0; d < 3; d++) vs[d] /= (NN - p->N_anchor);
  for (int i = 0; i < NN; i++)
    for (int d = 0; d < 3; d++) {
      double &v = vel[(size_t)d * NN + i];
      v -= vs[d];
      v2 += v * v;
    }
  for (int a = 0; a < p->N_anchor; a++)
    for (int d = 0; d < 3; d++) {
      double &v = vel[(size_t)d * NN + anchor[a] - 1];
      v2 -= v * v;
      v = 0;
    }
  v2 /= (NN - p->N_anchor - 1);
  const double s = std::sqrt(3.0 / p->Beta / v2);
  for (size_t k = 0; k < 3 * (size_t)NN; k++) vel[k] *= s;
  return MD_OK;
}
