// md_engine.cpp -- host side of the brush engine: device allocation, launch
// plans over many replicas, graph replay of the step loop, and the C-ABI of
// include/md_engine.h.
//
// An engine holds replicas (independent systems, possibly of different
// sizes, one precision).  Every kernel launch covers all replicas; a replica
// owns its device buffers, so adding or removing one never moves another's
// state.  The single-system entry points act on replica 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/md_engine.h"
#include "md_kernels.h"

using mdk::Ctrl;
using mdk::Plan;
using mdk::Rep;
using mdk::WMap;

namespace {

constexpr double kPi = 3.141592653589793;
// part, chg, chgG, fx, fy, fz, cell1, cell1w, cell2, cell3, cell3c, brick, plane, partG,
// rlst, cell2s
constexpr int kMaps = 16;
// LDS a workgroup may take on this device (both planes of a fused FFT
// plane kernel must fit)
size_t lds_limit() {
  static size_t lim = 0;
  if (!lim) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
      lim = (size_t)v;
    else
      lim = 65536;
  }
  return lim;
}

bool is_pow2(int n) { return n >= 4 && (n & (n - 1)) == 0; }
int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}
int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Cardinal B-spline M_p(x) by the Cox-de Boor recursion.
double bspline(double x, int p) {
  if (p == 2) return (x >= 0 && x <= 2) ? 1 - std::fabs(x - 1) : 0;
  return (x * bspline(x, p - 1) + (p - x) * bspline(x - 1, p - 1)) / (p - 1);
}

// |b(k)|^2 of the Euler exponential spline, pmeOrthoTabBC
// (compt_acc.f90:1636-1647), with its 1e15 guard for vanishing sums.
std::vector<double> bmod2(int K, int p) {
  std::vector<double> out(K);
  for (int k = 0; k < K; k++) {
    double re = 0, im = 0;
    for (int j = 0; j <= p - 2; j++) {
      const double ang = 2 * kPi * k * j / K;
      const double m = bspline(j + 1.0, p);
      re += m * std::cos(ang);
      im += m * std::sin(ang);
    }
    double mag2 = re * re + im * im;
    if (std::sqrt(mag2) < (double)1e-10f) mag2 = (double)1e15f * (double)1e15f;
    out[k] = 1.0 / mag2;
  }
  return out;
}

// Waits for the copies and fills issued on the null stream.  A hipMemcpy
// from pageable host memory, and a hipMemset, may return before the device
// has finished them, and the engine's stream (non-blocking) does not wait
// for the null stream: every such operation a kernel will read is followed
// by this before the kernel is launched.
hipError_t null_stream_done() { return hipStreamSynchronize(nullptr); }

// Read-only device tables (influence function, twiddles) shared by every
// replica with the same mesh, order, splitting and box: built once, freed
// with their last user.
struct DevTable {
  void *ptr = nullptr;
  ~DevTable() {
    if (ptr) (void)hipFree(ptr);
  }
};
std::mutex g_tables_mu;
std::map<std::string, std::weak_ptr<DevTable>> g_tables;

std::shared_ptr<DevTable> shared_table(const std::string &key,
                                       const std::function<std::vector<char>()> &build) {
  std::lock_guard<std::mutex> lock(g_tables_mu);
  int dev = 0;
  (void)hipGetDevice(&dev);
  const std::string k = std::to_string(dev) + "|" + key;
  auto it = g_tables.find(k);
  if (it != g_tables.end())
    if (auto t = it->second.lock()) return t;
  std::vector<char> host = build();
  auto t = std::make_shared<DevTable>();
  if (hipMalloc(&t->ptr, host.size()) != hipSuccess) {
    t->ptr = nullptr;
    return nullptr;
  }
  if (hipMemcpy(t->ptr, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess ||
      null_stream_done() != hipSuccess)
    return nullptr;
  g_tables[k] = t;
  return t;
}

std::string hip_msg(hipError_t e, const char *what) {
  return std::string(what) + ": " + hipGetErrorString(e);
}

// One replica: parameters, descriptor, owned device memory.
struct Sys {
  md_params p{};
  bool fp32 = false, coulomb = false, uploaded = false;
  bool real_list_wanted = false;  // md_build_lists asked for the reference-form real list
  Rep h{};
  std::vector<int> charge, flags, bstart, bpart;  // host topology
  std::vector<void *> allocs;
  std::vector<std::shared_ptr<DevTable>> tables;  // shared read-only tables
  size_t bytes = 0;
  int ncell1 = 0, ncell2 = 0, ncell3 = 0;
  int nb[kMaps] = {0};
  size_t lds_x = 0, lds_y = 0, lds_z = 0, lds_sort = 0, lds_sort3 = 0, lds_plane = 0;
  int nbin = 0;   // spreading tiles
  int nlists = 0; // spreading lists: one per tile or one per z index, whichever is more
  int nb10_wave = 0, nb10_wg = 0;  // real-space sweep grids of the two forms (the
                                   // workgroup form's before the engine's cap)
  int pz_big = 1;                  // z-pass lines per workgroup in large engines
  bool can_fuse = false;           // both planes of a fused FFT kernel fit in LDS
  size_t plane_bytes = 0;
  double *d_eta = nullptr, *d_out1 = nullptr, *d_out2 = nullptr;
  double4 *posb[2] = {nullptr, nullptr};  // the two position buffers (h.pos is posb[par])
  float4 *pos32b[2] = {nullptr, nullptr};
  int *d_hist = nullptr;     // observables: z histogram of the chains
  double *d_hsum = nullptr;  // observables: per-block sums of z
  double *d_mol = nullptr;   // observables: per-molecule heights
  double *d_epart = nullptr; // energies: per-block partial sums
  // dry run: allocations are only counted (no device call), for the bytes a
  // replica would take before it is admitted
  bool dry = false;
  int step_dev = -1;  // the control block's step counter when known on the host (md_run), else -1
  int en_blocks[3] = {0, 0, 0};  // energy kernels' blocks: particles, charges, mesh
  std::string err;

  ~Sys() {
    for (void *a : allocs) (void)hipFree(a);
  }
  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  template <typename T> T *dalloc(size_t n) {
    if (n == 0) n = 1;
    if (dry) {  // a placeholder address that is never dereferenced
      bytes += n * sizeof(T);
      return reinterpret_cast<T *>(alignof(double4));
    }
    void *ptr = nullptr;
    if (hipMalloc(&ptr, n * sizeof(T)) != hipSuccess) return nullptr;
    (void)hipMemset(ptr, 0, n * sizeof(T));
    allocs.push_back(ptr);
    bytes += n * sizeof(T);
    return (T *)ptr;
  }
  void dfree(void *ptr, size_t nbytes) {
    if (dry) {
      bytes -= nbytes;
      return;
    }
    auto it = std::find(allocs.begin(), allocs.end(), ptr);
    if (it != allocs.end()) {
      (void)hipFree(ptr);
      allocs.erase(it);
      bytes -= nbytes;
    }
  }
  int setup(const md_params *pin, bool want_fp32);
  int size_lists(const double *pos);
  // LJ cell and list capacities, and the walked real-space list's, as given
  int set_list_caps(int ccap1, int nbcap1, int nbcap2);
  // descriptor whose canonical positions are buffer q
  Rep table(int q) const {
    Rep t = h;
    t.pos = posb[q];
    t.pos32 = pos32b[q];
    t.pos_nx = posb[q ^ 1];
    t.pos32_nx = pos32b[q ^ 1];
    t.par = q;
    return t;
  }
  void set_par(int q) { h = table(q); }
};

int Sys::setup(const md_params *pin, bool want_fp32) {
  p = *pin;
  fp32 = want_fp32;
  coulomb = p.qq != 0 && p.Nq > 0;
  const int NN = p.NN, Nq = coulomb ? p.Nq : 0;
  if (NN <= 0 || p.Lx <= 0 || p.dt <= 0 || p.multistep <= 0) return fail(MD_EINVAL, "bad sizes");
  if (coulomb) {
    for (int d = 0; d < 3; d++) {
      if (p.ordr[d] < 3 || p.ordr[d] > mdk::kMaxOrder)
        return fail(MD_EINVAL, "SPME order must be in 3..8 on every axis");
      if (p.ng[d] < p.ordr[d] || p.ng[d] < 2) return fail(MD_EINVAL, "SPME mesh smaller than the spline order");
      if (p.ng[d] > (1 << 14)) return fail(MD_EINVAL, "SPME mesh larger than 16384 along an axis");
    }
  }
  // topology (compt_acc.f90:317-372, 433-488)
  std::vector<int32_t> chg(std::max(Nq, 1)), fene(2 * (size_t)std::max(p.N_bond, 1)),
      anchor(std::max(p.N_anchor, 1));
  int rc = md_build_topology(&p, coulomb ? chg.data() : nullptr, fene.data(), anchor.data());
  if (rc) return fail(rc, "topology");
  flags.assign(NN, 0);
  for (int a = 0; a < p.N_anchor; a++) flags[anchor[a] - 1] |= 1;
  charge.resize(Nq);
  for (int m = 0; m < Nq; m++) {
    charge[m] = chg[m] - 1;
    flags[chg[m] - 1] |= 2;
  }
  std::vector<int> deg(NN + 1, 0);
  for (int b = 0; b < p.N_bond; b++) {
    deg[fene[b] - 1]++;
    deg[fene[p.N_bond + b] - 1]++;
  }
  bstart.assign(NN + 1, 0);
  for (int i = 0; i < NN; i++) bstart[i + 1] = bstart[i] + deg[i];
  bpart.assign(std::max(bstart[NN], 1), 0);
  std::vector<int> fill(bstart.begin(), bstart.end() - 1);
  for (int b = 0; b < p.N_bond; b++) {
    const int j = fene[b] - 1, k = fene[p.N_bond + b] - 1;
    bpart[fill[j]++] = k;
    bpart[fill[k]++] = j;
  }

  // descriptor scalars, computed as the reference computes them
  Rep &S = h;
  S.NN = NN;
  S.Nq = Nq;
  S.Npe = p.Npe;
  S.N_anchor = p.N_anchor;
  S.real_verlet = coulomb ? p.real_verlet : 0;
  S.Lx = p.Lx;
  S.Ly = p.Ly;
  S.Lz = p.Lz;
  S.hLx = p.Lx / 2;
  S.hLy = p.Ly / 2;
  S.dt = p.dt;
  S.c1m = 1 - p.xi * p.dt / 2;
  S.c1p = 1 + p.xi * p.dt / 2.;
  S.sig = std::sqrt(2. / p.Beta / p.dt * p.xi);
  S.Beta = p.Beta;
  S.EF = p.EF;
  S.seed = p.seed;
  S.rcl2 = p.rcl * p.rcl;
  S.rcl3 = p.rcl_s * p.rcl_s;
  S.rcl4 = p.rcl_sp * p.rcl_sp;
  S.sigma2 = p.sigma_s * p.sigma_s;
  S.sigma3 = p.sigma_sp * p.sigma_sp;
  S.sigma_s = p.sigma_s;
  const double c5 = (double)(float)std::pow(5.0, (double)(1.0f / 6.0f));
  S.wall_r = p.rcl / c5;
  S.wall_rs = p.rcl_s / c5;
  S.wall_c = (double)0.4f;
  S.kR02 = p.kFENE * p.R0_2;
  S.R0_2 = p.R0_2;
  S.c1 = p.lb / p.Beta;
  S.alpha = p.alpha;
  S.alpha2 = p.alpha * p.alpha;
  S.c2 = 2 * p.alpha / std::sqrt(kPi);
  S.rcc2 = p.rcc * p.rcc;
  S.slab = 4 * kPi / p.Lx / p.Ly / p.Lz / p.Zfac;
  S.rv1sq = p.rvl_eff * p.rvl_eff;
  S.rv2sq = p.rvc * p.rvc;
  S.skin1 = (p.rvl_eff - p.rcl_s) / 2;
  S.skin2 = p.rsk / 2;
  const double L3[3] = {p.Lx, p.Ly, p.Lz};
  for (int d = 0; d < 3; d++) {
    S.nc1[d] = (int)(L3[d] / p.rvl_eff);
    S.rcel1[d] = S.nc1[d] > 0 ? L3[d] / S.nc1[d] : 0;
    S.nc2[d] = S.real_verlet ? (int)(L3[d] / p.rvc) : 1;
    S.rcel2[d] = S.real_verlet ? L3[d] / S.nc2[d] : L3[d];
  }
  if (S.nc1[0] < 1 || S.nc1[1] < 1 || S.nc1[2] < 1) return fail(MD_EINVAL, "box smaller than rvl");
  if (S.real_verlet && (S.nc2[0] < 1 || S.nc2[1] < 1 || S.nc2[2] < 1))
    return fail(MD_EINVAL, "box smaller than the real-space list radius");
  ncell1 = S.nc1[0] * S.nc1[1] * S.nc1[2];
  ncell2 = S.real_verlet ? S.nc2[0] * S.nc2[1] * S.nc2[2] : 0;
  // with fewer than three list cells along x or y the reference's stencil
  // lists a neighbour cell's pairs twice (or three times with one cell) and
  // its real-space sum counts every repeat: such replicas keep the
  // reference-form list, rebuilt at the reference's rebuild steps, and walk
  // it instead of sweeping cells
  S.real_list = S.real_verlet && (S.nc2[0] <= 2 || S.nc2[1] <= 2);
  if (S.real_list) real_list_wanted = true;
  // real-space sweep cells: edge >= r_c (one cell when the box is shorter);
  // with fewer than three cells along x or y the stencil visits each once
  S.real_cells = 0;
  ncell3 = 0;
  if (coulomb && !S.real_list) {
    for (int d = 0; d < 3; d++) {
      // cells of edge >= r_c (times the caller's cell scale, >= 1)
      const double cs = p.ovr_cell_scale > 1 ? p.ovr_cell_scale : 1.0;
      S.nc3[d] = std::max(1, (int)(L3[d] / (p.rcc * cs)));
      S.rcel3[d] = L3[d] / S.nc3[d];
    }
    S.real_cells = 1;
    ncell3 = S.nc3[0] * S.nc3[1] * S.nc3[2];
  }

  // state and topology on the device
  const size_t nn = NN;
  for (int q = 0; q < 2; q++) {
    posb[q] = dalloc<double4>(nn);
    pos32b[q] = fp32 ? dalloc<float4>(nn) : nullptr;
    if (!posb[q] || (fp32 && !pos32b[q])) return fail(MD_ENOMEM, "state");
  }
  set_par(0);
  double *v = dalloc<double>(3 * nn), *a = dalloc<double>(3 * nn);
  S.ngrp = cdiv(NN, mdk::kSRParticles);  // the short-range kernel's particle groups
  S.dpart = dalloc<double2>(2 * (size_t)S.ngrp);
  if (!v || !a || !S.dpart) return fail(MD_ENOMEM, "state");
  S.vx = v; S.vy = v + nn; S.vz = v + 2 * nn;
  S.ax = a; S.ay = a + nn; S.az = a + 2 * nn;
  const size_t rs = fp32 ? 4 : 8;
  char *cc = dalloc<char>(3 * nn * rs);
  S.ccx = cc; S.ccy = cc + nn * rs; S.ccz = cc + 2 * nn * rs;
  S.flags = dalloc<int>(nn);
  S.charge = dalloc<int>(std::max(Nq, 1));
  S.bstart = dalloc<int>(nn + 1);
  S.bpart = dalloc<int>(bpart.size());
  S.bond4 = dalloc<int4>(nn);
  S.ctrl = dalloc<Ctrl>(1);
  d_eta = dalloc<double>(3 * nn);
  d_out1 = dalloc<double>(3 * nn);
  d_out2 = dalloc<double>(3 * nn);
  d_hist = dalloc<int>(mdk::kHistBins);
  d_hsum = dalloc<double>(cdiv(NN, mdk::kBlock));  // one per particle block of k_observe
  d_mol = dalloc<double>(4 * (size_t)std::max(1, p.Nga + p.Ngl));  // per-molecule heights
  en_blocks[0] = cdiv(NN, mdk::kBlock);
  en_blocks[1] = coulomb ? cdiv(Nq, mdk::kBlock) : 0;
  en_blocks[2] = coulomb ? (int)std::min<long>(256, cdiv((long)p.ng[0] * p.ng[1] * p.ng[2], mdk::kBlock)) : 0;
  d_epart = dalloc<double>((size_t)en_blocks[0] * mdk::kEnShortTerms + 2 * (size_t)en_blocks[1] + en_blocks[2]);
  S.Nga = p.Nga;
  S.arm = p.arm;
  S.Nma = p.Nma;
  S.Ngl = p.Ngl;
  S.Nml = p.Nml;
  if (!cc || !S.flags || !S.charge || !S.bstart || !S.bpart || !S.bond4 || !S.ctrl || !d_eta || !d_out1 ||
      !d_out2 || !d_hist || !d_hsum || !d_mol || !d_epart)
    return fail(MD_ENOMEM, "topology arrays");
  hipError_t cerr = dry ? hipSuccess : hipMemcpy(S.flags, flags.data(), nn * sizeof(int), hipMemcpyHostToDevice);
  if (cerr == hipSuccess && Nq && !dry)
    cerr = hipMemcpy(S.charge, charge.data(), Nq * sizeof(int), hipMemcpyHostToDevice);
  if (cerr == hipSuccess && !dry)
    cerr = hipMemcpy(S.bstart, bstart.data(), (nn + 1) * sizeof(int), hipMemcpyHostToDevice);
  if (cerr == hipSuccess && !dry)
    cerr = hipMemcpy(S.bpart, bpart.data(), bpart.size() * sizeof(int), hipMemcpyHostToDevice);
  if (cerr == hipSuccess && !dry) {
    std::vector<int4> b4(nn);
    for (int i = 0; i < NN; i++) {
      int v[4] = {-1, -1, -1, -1};
      for (int k = 0; k < 4 && bstart[i] + k < bstart[i + 1]; k++) v[k] = bpart[bstart[i] + k];
      b4[i] = make_int4(v[0], v[1], v[2], v[3]);
    }
    cerr = hipMemcpy(S.bond4, b4.data(), nn * sizeof(int4), hipMemcpyHostToDevice);
  }
  if (cerr != hipSuccess) return fail(MD_EHIP, hip_msg(cerr, "topology upload"));
  S.cnt1 = dalloc<int>(2 * (size_t)ncell1);  // two count buffers, used in turn
  S.cell1_of = dalloc<int>(NN);
  if (S.cell1_of && !dry) (void)hipMemset(S.cell1_of, 0xff, NN * sizeof(int));  // no cell yet
  S.cnt2[0] = dalloc<int>(std::max(ncell2, 1));
  S.cnt2[1] = dalloc<int>(std::max(ncell2, 1));
  S.cnt3 = dalloc<int>(std::max(ncell3, 1));
  S.scnt3 = dalloc<int>(std::max(ncell3, 1));
  S.nbn1 = dalloc<int>(nn);
  S.nbn2 = dalloc<int>(std::max(Nq, 1));
  if (!S.cnt1 || !S.cell1_of || !S.cnt2[0] || !S.cnt2[1] || !S.cnt3 || !S.scnt3 || !S.nbn1 ||
      !S.nbn2)
    return fail(MD_ENOMEM, "cells");

  // SPME: mesh, spectrum, influence function, twiddles
  const int csz = fp32 ? 8 : 16;
  if (coulomb) {
    const int K1 = p.ng[0], K2 = p.ng[1], K3 = p.ng[2];
    for (int d = 0; d < 3; d++) {
      S.K[d] = p.ng[d];
      S.pw2[d] = is_pow2(p.ng[d]);
      S.lK[d] = S.pw2[d] ? ilog2(p.ng[d]) : 0;
      // other lengths: passes of radix 8, 4, 2 for the power-of-two part,
      // then 5, 3 and the remaining primes
      S.nfac[d] = 0;
      if (!S.pw2[d]) {
        int n = p.ng[d];
        auto push = [&](int f) { S.fac[d][S.nfac[d]++] = f; };
        while (n % 8 == 0) { push(8); n /= 8; }
        while (n % 4 == 0) { push(4); n /= 4; }
        while (n % 2 == 0) { push(2); n /= 2; }
        while (n % 5 == 0) { push(5); n /= 5; }
        while (n % 3 == 0) { push(3); n /= 3; }
        for (int f = 7; n > 1; f += 2)
          while (n % f == 0) { push(f); n /= f; }
      }
      S.kg[d] = p.ng[d] / p.gdim[d];
      S.fnorm[d] = p.ng[d] / p.gdim[d] / ((double)K1 * K2 * K3);
      S.ord[d] = p.ordr[d];
    }
    S.order = (p.ordr[1] == p.ordr[0] && p.ordr[2] == p.ordr[0]) ? p.ordr[0] : 0;
    S.Kh = K1 / 2 + 1;

    const size_t M = (size_t)K1 * K2 * K3, Mh = (size_t)K3 * K2 * S.Kh;
    S.meshr = dalloc<char>(M * rs);
    S.spec = dalloc<char>(Mh * csz);
    S.meshu = dalloc<char>(M * rs);
    if (!S.meshr || !S.spec || !S.meshu) return fail(MD_ENOMEM, "mesh");
    // spreading tiles: ceil(K / edge) per axis of near-equal edges (K / ntl
    // or one more), edge kTile or the caller's (4..kTile), every edge >=
    // the axis's order (fewer tiles where it would not be), so a charge's
    // points fall in at most two tiles along each axis
    nbin = 1;
    for (int d = 0; d < 3; d++) {
      const int edge = p.ovr_tile[d] > 0 ? std::min(std::max(p.ovr_tile[d], 4), mdk::kTile) : mdk::kTile;
      S.ntl[d] = cdiv(p.ng[d], edge);
      if (p.ng[d] / S.ntl[d] < S.ord[d]) S.ntl[d] = std::max(1, p.ng[d] / S.ord[d]);
      S.tsz[d] = p.ng[d] / S.ntl[d];
      S.trem[d] = p.ng[d] % S.ntl[d];
      nbin *= S.ntl[d];
    }
    // the same lists serve plane spreading, one per first z index
    nlists = std::max(nbin, p.ng[2]);
    S.nlist = nlists * mdk::kTileSub;
    S.tcnt = dalloc<int>((size_t)nlists * mdk::kTileSub);
    S.tmem = nullptr;
    S.capt = 0;
    if (!S.tcnt) return fail(MD_ENOMEM, "spreading tiles");
    // x and y passes fused per z-plane when both planes fit in LDS
    // two regions of max(K2/2 * K1, Kh * K2) points (the FFT passes alternate)
    // (plus the x and y twiddle tables, staged in LDS, and plane
    // spreading's sub-list table)
    plane_bytes = (2 * mdk::plane_region(K1, K2, S.Kh) + K1 + K2) * csz + mdk::kPlaneSpreadTable;
    lds_plane = plane_bytes;
    // MD_FFT_PLANES=0 keeps the separate x and y passes (for comparison)
    const char *fp = std::getenv("MD_FFT_PLANES");
    // (at most 64 KB: larger planes leave one workgroup per CU and lose)
    can_fuse = lds_plane <= std::min<size_t>(65536, dry ? 65536 : lds_limit()) && !(fp && fp[0] == '0');
    S.plane_fused = can_fuse;
    if (!S.plane_fused) lds_plane = 0;
    // influence function of PME_BC (compt_acc.f90:1557-1598), half spectrum
    // (shared between replicas and not counted in a replica's bytes: a dry
    // run skips them)
    if (!dry) {
    char key[256];
    std::snprintf(key, sizeof key, "bc %d %d %d %d %d %d %a %a %a %a %d %d", K1, K2, K3, p.ordr[0],
                  p.ordr[1], p.ordr[2], p.alpha, p.gdim[0], p.gdim[1], p.gdim[2], (int)fp32, 0);
    auto bc = shared_table(key, [&]() {
      std::vector<double> b2[3], mv[3], ev[3];
      for (int d = 0; d < 3; d++) {
        const int K = p.ng[d];
        b2[d] = bmod2(K, p.ordr[d]);
        mv[d].resize(K);
        ev[d].resize(K);
        for (int k = 0; k < K; k++) {
          mv[d][k] = (k <= K / 2 ? k : k - K) / p.gdim[d];
          const double t = kPi / p.alpha;
          ev[d][k] = std::exp(-t * t * mv[d][k] * mv[d][k]);
        }
      }
      const double pivol = (double)K1 * K2 * K3 / (kPi * p.gdim[0] * p.gdim[1] * p.gdim[2]);
      std::vector<char> out(Mh * rs);
      for (int z = 0; z < K3; z++)
        for (int y = 0; y < K2; y++) {
          const int ky = y;
          for (int x = 0; x < S.Kh; x++) {
            const double m2 = mv[0][x] * mv[0][x] + mv[1][ky] * mv[1][ky] + mv[2][z] * mv[2][z];
            const double v = (x + ky + z == 0) ? 0.0
                                                : b2[0][x] * b2[1][ky] * b2[2][z] * pivol * ev[0][x] *
                                                      ev[1][ky] * ev[2][z] / m2;
            const size_t at = ((size_t)z * K2 + y) * S.Kh + x;
            if (fp32) ((float *)out.data())[at] = (float)v;
            else ((double *)out.data())[at] = v;
          }
        }
      return out;
    });
    if (!bc) return fail(MD_ENOMEM, "influence function");
    tables.push_back(bc);
    S.bc = bc->ptr;
    for (int d = 0; d < 3; d++) {
      const int K = p.ng[d];
      std::snprintf(key, sizeof key, "tw %d %d", K, (int)fp32);
      auto tw = shared_table(key, [&]() {
        std::vector<char> out((size_t)K * csz);  // K complex: the full circle
        for (int k = 0; k < K; k++) {
          const double c = std::cos(2 * kPi * k / K), sn = -std::sin(2 * kPi * k / K);
          if (fp32) {
            ((float *)out.data())[2 * k] = (float)c;
            ((float *)out.data())[2 * k + 1] = (float)sn;
          } else {
            ((double *)out.data())[2 * k] = c;
            ((double *)out.data())[2 * k + 1] = sn;
          }
        }
        return out;
      });
      if (!tw) return fail(MD_ENOMEM, "twiddles");
      tables.push_back(tw);
      S.tw[d] = tw->ptr;
    }
    }
    // FFT tiles: about 32 KB of LDS per workgroup
    const int budget = 16384;  // bytes per LDS buffer (the FFT passes use two)
    S.px_tile = std::max(1, std::min((K2 * K3 + 1) / 2, budget / (K1 * csz)));
    S.py_tile = std::max(1, std::min(S.Kh, budget / (K2 * csz)));
    // z pass lines per workgroup: set at rebuild with the engine's size
    S.pz_tile = std::max(1, std::min(4, budget / (K3 * csz)));
    pz_big = std::max(1, std::min(16, budget / (K3 * csz)));
    S.nqblk = cdiv((long)Nq, mdk::kBlock / mdk::kGatherLanes);  // gather blocks
    // per particle (the charged ones written), so that the short-range
    // kernel reads its particles' entries without the charge slot
    S.fsp = dalloc<double>(3 * (size_t)NN);
    S.freal = dalloc<char>(3 * (size_t)NN * (fp32 ? 4 : 8));
    S.part = dalloc<double>(4 * (size_t)S.nqblk);
    if (!S.fsp || !S.part || !S.freal) return fail(MD_ENOMEM, "spme scratch");
    lds_x = (2 * (size_t)S.px_tile * K1 + K1) * csz;
    lds_y = (2 * (size_t)S.py_tile * mdk::plane_line_stride(K2) + K2) * csz;
    lds_z = (2 * (size_t)S.pz_tile * mdk::zline_stride(K3) + K3) * csz;
  }
  nb[0] = cdiv(NN, mdk::kBlock);
  nb[1] = coulomb ? cdiv(Nq, mdk::kBlock) : 0;
  nb[2] = coulomb ? S.nqblk : 0;
  nb[3] = coulomb && !S.plane_fused ? cdiv(((long)S.K[1] * S.K[2] + 1) / 2, S.px_tile) : 0;
  nb[4] = coulomb && !S.plane_fused ? S.K[2] * cdiv(S.Kh, S.py_tile) : 0;
  nb[5] = coulomb ? cdiv((long)S.K[1] * S.Kh, S.pz_tile) : 0;
  // cell kernels stride over the cells with a grid of modest size
  nb[6] = nb[7] = 0;  // the LJ list kernels use the rebuild queue
  nb[8] = ncell2;
  nb[9] = ncell3;
  nb[10] = 0;  // set with the cell capacity (size_lists)
  nb[11] = coulomb ? nbin : 0;
  nb[12] = coulomb && S.plane_fused ? S.K[2] : 0;
  nb[13] = cdiv(NN, mdk::kSRParticles);
  nb[14] = S.real_list ? cdiv(Nq, mdk::kBlock / 64) : 0;  // a wavefront per charge
  nb[15] = S.real_list ? ncell2 : 0;
  S.mem1 = S.mem2 = S.mem3 = nullptr;
  S.pos3s = nullptr;
  S.nb1 = S.nb2 = nullptr;
  S.ccap1 = S.ccap2 = S.ccap3 = S.nbcap1 = S.nbcap2 = 0;
  if (!dry && null_stream_done() != hipSuccess) return fail(MD_EHIP, "set-up copies");
  return MD_OK;
}

// Sizes cell and list capacities from a configuration: twice the largest
// occupancy and list length found on the host, plus headroom.
int Sys::set_list_caps(int ccap1, int nbcap1, int nbcap2) {
  if (ccap1 <= 0 || nbcap1 <= 0) return fail(MD_EINVAL, "list capacities");
  if (h.real_list && nbcap2 != h.nbcap2) {
    if (h.nb2) dfree(h.nb2, (size_t)h.Nq * h.nbcap2 * 4);
    h.nb2 = dalloc<int>((size_t)h.Nq * nbcap2);
    h.nbcap2 = nbcap2;
    if (!h.nb2) return fail(MD_ENOMEM, "real-space list");
  }
  if (ccap1 != h.ccap1) {
    if (h.mem1) dfree(h.mem1, (size_t)ncell1 * h.ccap1 * 4);
    h.mem1 = dalloc<int>((size_t)ncell1 * ccap1);
    h.ccap1 = ccap1;
  }
  if (nbcap1 != h.nbcap1) {
    if (h.nb1) dfree(h.nb1, (size_t)p.NN * h.nbcap1 * 4);
    h.nb1 = dalloc<int>((size_t)p.NN * nbcap1);
    h.nbcap1 = nbcap1;
  }
  if (!h.mem1 || !h.nb1) return fail(MD_ENOMEM, "neighbour lists");
  if (null_stream_done() != hipSuccess) return fail(MD_EHIP, "list buffers");
  return MD_OK;
}

int Sys::size_lists(const double *pos) {
  const int NN = p.NN;
  const double *X = pos, *Y = pos + NN, *Z = pos + 2 * (size_t)NN;
  double occ3 = 0;  // mean occupancy of a charge's real-space cell
  auto scan = [&](const std::vector<int> &items, const int nc[3], double rv, int &maxcell,
                  int &maxnb) {
    const double rc[3] = {p.Lx / nc[0], p.Ly / nc[1], p.Lz / nc[2]};
    const int ncell = nc[0] * nc[1] * nc[2];
    std::vector<std::vector<int>> cells(ncell);
    std::vector<int> cid(items.size());
    for (size_t t = 0; t < items.size(); t++) {
      const int i = items[t];
      int a = (int)((X[i] + p.Lx / 2) / rc[0]), b = (int)((Y[i] + p.Ly / 2) / rc[1]),
          c = (int)(Z[i] / rc[2]);
      a = std::min(std::max(a, 0), nc[0] - 1);
      b = std::min(std::max(b, 0), nc[1] - 1);
      c = std::min(std::max(c, 0), nc[2] - 1);
      cid[t] = (c * nc[1] + b) * nc[0] + a;
      cells[cid[t]].push_back(i);
    }
    maxcell = 0;
    double s2 = 0;
    for (auto &c : cells) {
      maxcell = std::max(maxcell, (int)c.size());
      s2 += (double)c.size() * c.size();
    }
    occ3 = items.empty() ? 0 : s2 / items.size();
    maxnb = 0;
    if (rv <= 0) return;  // occupancy only
    for (size_t t = 0; t < items.size(); t++) {
      const int i = items[t], cell = cid[t];
      const int a = cell % nc[0], b = (cell / nc[0]) % nc[1], c = cell / (nc[0] * nc[1]);
      int n = 0;
      for (int l = -1; l <= 1; l++)
        for (int m = -1; m <= 1; m++)
          for (int k = -1; k <= 1; k++) {
            const int z = c + k;
            if (z < 0 || z >= nc[2]) continue;
            const int xx = (a + l + nc[0]) % nc[0], yy = (b + m + nc[1]) % nc[1];
            for (int j : cells[(z * nc[1] + yy) * nc[0] + xx]) {
              if (j == i) continue;
              double dx = X[i] - X[j], dy = Y[i] - Y[j], dz = Z[i] - Z[j];
              if (dx > p.Lx / 2) dx -= p.Lx; else if (dx <= -p.Lx / 2) dx += p.Lx;
              if (dy > p.Ly / 2) dy -= p.Ly; else if (dy <= -p.Ly / 2) dy += p.Ly;
              if (dx * dx + dy * dy + dz * dz < rv * rv) n++;
            }
          }
      maxnb = std::max(maxnb, n);
    }
  };
  std::vector<int> all(NN);
  for (int i = 0; i < NN; i++) all[i] = i;
  int mc1, mn1, mc2 = 0, mn2 = 0, mc3 = 0, mn3 = 0;
  scan(all, h.nc1, p.rvl_eff, mc1, mn1);
  // the reference-form real-space list is built only on request
  // (md_build_lists): its neighbour scan, the costly part of an upload, waits
  // for the first request
  if (h.real_verlet && real_list_wanted) scan(charge, h.nc2, p.rvc, mc2, mn2);
  if (h.real_cells) {
    scan(charge, h.nc3, 0.0, mc3, mn3);
    (void)occ3;
  }
  auto round8 = [](int v) { return (v + 7) & ~7; };
  // LJ: dense packing bounds occupancy near 1.4 per unit volume
  const double v1 = h.rcel1[0] * h.rcel1[1] * h.rcel1[2];
  // at most 128: the fill pass ranks a cell's members by counting
  const int ccap1 = std::min(128, round8(std::max(2 * mc1 + 16, (int)(1.5 * v1) + 16)));
  const int nbcap1 =
      round8(std::min(NN, std::max(2 * mn1 + 32, (int)(4.19 * std::pow(p.rvl_eff, 3) * 1.2) + 32)));
  const int ccap2 = h.real_verlet ? round8(std::min(h.Nq, 2 * mc2 + 64)) : 0;
  const int nbcap2 = h.real_verlet ? round8(std::min(h.Nq, 2 * mn2 + 64)) : 0;
  const int ccap3 = h.real_cells ? round8(std::min(h.Nq, 2 * mc3 + 64)) : 0;
  if (ccap2 > 16384 || ccap3 > 16384)
    return fail(MD_EINVAL, "real-space cells too full for the LDS sort");
  if (h.real_cells && ccap3 != h.ccap3) {
    const size_t vsz = fp32 ? 16 : 32;
    if (h.mem3) dfree(h.mem3, (size_t)ncell3 * h.ccap3 * 4);
    if (h.pos3s) dfree(h.pos3s, (size_t)ncell3 * h.ccap3 * vsz);
    h.mem3 = dalloc<int>((size_t)ncell3 * ccap3);
    h.pos3s = dalloc<char>((size_t)ncell3 * ccap3 * vsz);
    if (!h.pos3s) return fail(MD_ENOMEM, "real-space cells");
    h.ccap3 = ccap3;
    if (!h.mem3) return fail(MD_ENOMEM, "real-space cells");
  }
  if (coulomb) {  // spreading tile lists
    std::vector<int> cnt((size_t)nbin * mdk::kTileSub, 0);
    std::vector<int> cntp((size_t)h.K[2] * mdk::kTileSub, 0);  // plane lists: first z index
    int mb = 0;
    for (int m = 0; m < (int)charge.size(); m++) {
      const int i = charge[m];
      const int sub = (m >> 3) & (mdk::kTileSub - 1);
      const double c3[3] = {X[i], Y[i], Z[i]};
      int t0[3], nt[3];
      for (int d = 0; d < 3; d++) {
        int b = ((int)std::ceil(h.kg[d] * c3[d]) - h.ord[d]) % h.K[d];
        b += b < 0 ? h.K[d] : 0;
        t0[d] = mdk::tile_of(h, d, b);
        nt[d] = mdk::tile_of(h, d, (b + h.ord[d] - 1) % h.K[d]) != t0[d] ? 2 : 1;
        if (d == 2) mb = std::max(mb, ++cntp[(size_t)b * mdk::kTileSub + sub]);
      }
      for (int c = 0; c < nt[2]; c++)
        for (int b = 0; b < nt[1]; b++)
          for (int a = 0; a < nt[0]; a++) {
            const int t = (((t0[2] + c) % h.ntl[2]) * h.ntl[1] + (t0[1] + b) % h.ntl[1]) * h.ntl[0] +
                          (t0[0] + a) % h.ntl[0];
            mb = std::max(mb, ++cnt[(size_t)t * mdk::kTileSub + sub]);
          }
    }
    const int capt = round8(std::min(h.Nq, 2 * mb + 32));
    if (capt != h.capt) {
      if (h.tmem) dfree(h.tmem, (size_t)nlists * mdk::kTileSub * h.capt * 4);
      h.tmem = dalloc<int>((size_t)nlists * mdk::kTileSub * capt);
      h.capt = capt;
      if (!h.tmem) return fail(MD_ENOMEM, "spreading tiles");
    }
  }
  int np3 = 1;
  while (np3 < h.ccap3) np3 <<= 1;
  lds_sort3 = h.real_cells ? (size_t)np3 * sizeof(int) : 0;
  // about one non-empty item per wavefront, or per workgroup (the form is
  // chosen for the whole engine at rebuild)
  nb10_wg = h.real_cells ? std::max(1, cdiv(h.Nq, mdk::kRealChunkWG)) : 0;
  nb10_wave =
      h.real_cells ? std::max(1, std::min(cdiv(h.Nq, mdk::kRealChunk * (mdk::kBlock / 64)), 4096)) : 0;
  nb[10] = h.real_wg ? std::min(nb10_wg, 1024) : nb10_wave;  // the engine's rebuild sets its cap
  if (ccap1 != h.ccap1) {
    if (h.mem1) dfree(h.mem1, (size_t)ncell1 * h.ccap1 * 4);
    h.mem1 = dalloc<int>((size_t)ncell1 * ccap1);
    h.ccap1 = ccap1;
  }
  if (nbcap1 != h.nbcap1) {
    if (h.nb1) dfree(h.nb1, (size_t)NN * h.nbcap1 * 4);
    h.nb1 = dalloc<int>((size_t)NN * nbcap1);
    h.nbcap1 = nbcap1;
  }
  if (h.real_verlet && ccap2 != h.ccap2) {
    if (h.mem2) dfree(h.mem2, (size_t)ncell2 * h.ccap2 * 4);
    h.mem2 = dalloc<int>((size_t)ncell2 * ccap2);
    h.ccap2 = ccap2;
  }
  if (h.real_verlet && nbcap2 != h.nbcap2) {
    if (h.nb2) dfree(h.nb2, (size_t)h.Nq * h.nbcap2 * 4);
    h.nb2 = dalloc<int>((size_t)h.Nq * nbcap2);
    h.nbcap2 = nbcap2;
  }
  if (!h.mem1 || !h.nb1 || (h.real_verlet && (!h.mem2 || !h.nb2)))
    return fail(MD_ENOMEM, "neighbour lists");
  int np = 1;
  while (np < h.ccap2) np <<= 1;
  lds_sort = h.real_verlet ? (size_t)np * sizeof(int) : 0;
  if (!dry && null_stream_done() != hipSuccess) return fail(MD_EHIP, "list buffers");
  return MD_OK;
}

}  // namespace

struct md_engine {
  bool fp32 = false;
  int multistep = 0;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;       // the real-space sweep of a Coulomb evaluation
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<Sys *> sys;          // launch order
  std::map<int, Sys *> slots;      // stable handles
  int next_slot = 0;
  // Replayed steps alternate between two position buffers: a step reads
  // buffer q and its short-range kernel drifts the next step's positions
  // into buffer q ^ 1.  Descriptor table q (and plans planp[q], singlep[q])
  // hold buffer q as canonical; cur is the buffer of the state now.
  int cur = 0;
  Plan planp[2];
  std::vector<Plan> singlep[2];
  Plan plan;                       // all replicas, buffer cur
  std::vector<Plan> single;        // one replica each (same order as sys), buffer cur
  Rep *d_reps = nullptr;           // table 0, then table 1, reps_cap each
  int reps_cap = 0;
  int *d_maps = nullptr;
  int maps_cap = 0;
  unsigned long long *d_lq = nullptr;  // LJ rebuild queues: the batch's, then one per replica
  size_t lq_words = 0;
  double *vpart = nullptr;
  size_t vpart_cap = 0;
  int *h_steps = nullptr;          // pinned, one per replica
  Ctrl *h_ctrl = nullptr;          // pinned, one per replica: control blocks read back
  int steps_cap = 0;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // cycle graphs by the buffer of their first step
  hipGraph_t graph[2] = {nullptr, nullptr};
  bool dirty = true;
  bool batch = false;              // made by md_batch_create (replicas pre-heat)
  // launch forms set by md_autotune (-1 / 0: the size heuristics); none
  // of them changes a result
  int tune_real = -1;              // real-space sweep: 0 wavefront, 1 workgroup per item
  int tune_pz = 0;                 // z-pass lines per workgroup
  int tune_fuse = -1;              // fused FFT planes where they fit: 0 no, 1 yes
  int tune_ride = -1;              // real-space sweep inside the spreading / gather launches
  int tune_rgrid = 1024;           // workgroups of the real-space sweep's workgroup form (cap)
  int tune_pspread = 1;            // spreading inside the forward plane kernel: 0 no, 1 yes
  int tune_rparts = 3;             // with plane spreading: launches the real-space sweep rides in
  int tune_lead = 0;               // riding sweep workgroups dispatched ahead of the plane launch's own
  int tune_lead_g = 0;             // ... ahead of the gather launch's own
  std::string err;

  ~md_engine() {
    if (st) (void)hipStreamSynchronize(st);
    drop_graph();
    for (Sys *s : sys) delete s;
    if (d_reps) (void)hipFree(d_reps);
    if (d_maps) (void)hipFree(d_maps);
    if (d_lq) (void)hipFree(d_lq);
    if (vpart) (void)hipFree(vpart);
    if (h_steps) (void)hipHostFree(h_steps);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    if (st2) (void)hipStreamSynchronize(st2);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (st2) (void)hipStreamDestroy(st2);
    if (st) (void)hipStreamDestroy(st);
  }
  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
  int hip_fail(hipError_t e, const char *what) { return fail(MD_EHIP, hip_msg(e, what)); }
  void drop_graph() {
    for (int q = 0; q < 2; q++) {
      if (gexec[q]) (void)hipGraphExecDestroy(gexec[q]);
      if (graph[q]) (void)hipGraphDestroy(graph[q]);
      gexec[q] = nullptr;
      graph[q] = nullptr;
    }
  }
  // makes buffer q the canonical one (host descriptors and current plans)
  void select(int q) {
    cur = q;
    for (Sys *x : sys) x->set_par(q);
    plan = planp[q];
    single = singlep[q];
  }
  int init(int precision, int ms) {
    fp32 = precision == MD_FP32;
    multistep = ms;
    if (ms <= 0) return fail(MD_EINVAL, "multistep must be positive");
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_join, hipEventDisableTiming);
    return e == hipSuccess ? MD_OK : hip_fail(e, "stream");
  }
  int rebuild();
  int sync();
  int check_one(const Sys *s, const Ctrl &c, std::string &msg) const;
  int check_all();
  int check_sys(const Sys *s);  // waits for the stream; the status of one replica
  int add(const md_params *p, int *slot);
  int launch_compute_force(const Plan &P, Sys *s, int step, int noise, const double *eta, bool in_step);
  int error_analysis(Sys *s, const Plan &P, double *out, double *acc_se);
  int error_analysis_noalloc(Sys *s, const Plan &P, double *out, double *acc_se);
  int build_graph(int q);
  int run_cycles(int ncycles);
  bool any_real_list() const {
    for (const Sys *x : sys)
      if (x->h.real_list) return true;
    return false;
  }
  int index_of(const Sys *s) const {
    for (size_t k = 0; k < sys.size(); k++)
      if (sys[k] == s) return (int)k;
    return -1;
  }
};

// Uploads every descriptor and the block maps; per-replica plans point into
// the same arrays.  Invalidates the captured graph.
int md_engine::rebuild() {
  const int R = (int)sys.size();
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "sync");
  drop_graph();
  dirty = false;
  if (R == 0) {
    planp[0] = planp[1] = Plan{};
    singlep[0].clear();
    singlep[1].clear();
    select(cur);
    return MD_OK;
  }
  if (R > reps_cap) {
    if (d_reps) (void)hipFree(d_reps);
    reps_cap = std::max(2 * R, 8);
    if (hipMalloc(&d_reps, 2 * (size_t)reps_cap * sizeof(Rep)) != hipSuccess)
      return fail(MD_ENOMEM, "descriptors");
  }
  // real-space sweep form: a workgroup per item (measured faster from one
  // default brush to the 256-replica serving batch) unless tuned; both forms
  // give the same sums (MD_REAL_SWEEP=wave|wg forces one)
  int wg = tune_real >= 0 ? tune_real : 1;
  const char *rsw = std::getenv("MD_REAL_SWEEP");
  if (rsw && !std::strcmp(rsw, "wave")) wg = 0;
  if (rsw && !std::strcmp(rsw, "wg")) wg = 1;
  // the workgroup form's grid: about one non-empty item per workgroup, at
  // most tune_rgrid workgroups (MD_REAL_GRID forces a cap); the items a
  // workgroup takes do not change any sum
  const char *rg = std::getenv("MD_REAL_GRID");
  const int rcap = rg ? std::max(1, std::atoi(rg)) : tune_rgrid;
  for (Sys *x : sys) {
    x->h.real_wg = wg;
    x->nb[10] = wg ? std::min(x->nb10_wg, rcap) : x->nb10_wave;  // 0 where the list is walked
  }
  const bool reps_host_wg = wg != 0;  // one form for every replica of the engine
  // z pass: few lines per workgroup (32-byte row pieces) while the engine's
  // mesh is small, so that the launch fills the chip; whole 128-byte pieces
  // when the engine holds many mesh points.  The lines' arithmetic is the
  // same either way.
  long mesh_all = 0;
  for (Sys *x : sys)
    if (x->coulomb) mesh_all += (long)x->h.K[0] * x->h.K[1] * x->h.K[2];
  for (Sys *x : sys) {
    if (!x->coulomb) continue;
    const int csz = x->fp32 ? 8 : 16;
    const int small = std::max(1, std::min(4, 16384 / (x->h.K[2] * csz)));
    x->h.pz_tile = mesh_all > (16L << 20) ? x->pz_big : small;
    const char *zl = std::getenv("MD_Z_LINES");  // forces the z-pass lines, for comparison
    const int want_pz = zl ? std::atoi(zl) : tune_pz;
    if (want_pz > 0) x->h.pz_tile = std::max(1, std::min(want_pz, 16384 / (x->h.K[2] * csz)));
    x->nb[5] = cdiv((long)x->h.K[1] * x->h.Kh, x->h.pz_tile);
    x->lds_z = (2 * (size_t)x->h.pz_tile * mdk::zline_stride(x->h.K[2]) + x->h.K[2]) * csz;
    // fused planes where they fit, unless tuned off
    Rep &S = x->h;
    const char *fp = std::getenv("MD_FFT_PLANES");  // =0 keeps separate x and y passes
    S.plane_fused = x->can_fuse && tune_fuse != 0 && !(fp && fp[0] == '0');
    x->lds_plane = S.plane_fused ? x->plane_bytes : 0;
    x->nb[3] = !S.plane_fused ? cdiv(((long)S.K[1] * S.K[2] + 1) / 2, S.px_tile) : 0;
    x->nb[4] = !S.plane_fused ? S.K[2] * cdiv(S.Kh, S.py_tile) : 0;
    x->nb[12] = S.plane_fused ? S.K[2] : 0;
  }
  // spreading inside the forward plane kernel: when every replica with
  // Coulomb forces has fused planes of at most kPlaneSpreadPer points per
  // lane, and it is asked for (tuned; MD_PLANE_SPREAD=0/1 forces); the
  // mesh is the same either way
  bool ps_ok = false;
  for (Sys *x : sys) {
    if (!x->coulomb) continue;
    ps_ok = true;
    if (!x->h.plane_fused || x->h.K[0] * x->h.K[1] > mdk::kPlaneSpreadPer * mdk::kPlaneBlock) {
      ps_ok = false;
      break;
    }
  }
  const char *ps = std::getenv("MD_PLANE_SPREAD");
  const bool pspread = ps_ok && (ps ? ps[0] == '1' : tune_pspread > 0);
  std::vector<Rep> hr(R);
  for (int q = 0; q < 2; q++) {
    for (int r = 0; r < R; r++) hr[r] = sys[r]->table(q);
    e = hipMemcpy(d_reps + (size_t)q * reps_cap, hr.data(), R * sizeof(Rep), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "descriptors");
  }
  // maps: batch maps kMaps*(R+1), then per replica kMaps*2
  std::vector<int> m;
  int tot[kMaps];
  for (int k = 0; k < kMaps; k++) {
    int acc = 0;
    m.push_back(0);
    for (int r = 0; r < R; r++) {
      acc += sys[r]->nb[k];
      m.push_back(acc);
    }
    tot[k] = acc;
  }
  const int single_off = (int)m.size();
  for (int r = 0; r < R; r++)
    for (int k = 0; k < kMaps; k++) {
      m.push_back(0);
      m.push_back(sys[r]->nb[k]);
    }
  if ((int)m.size() > maps_cap) {
    if (d_maps) (void)hipFree(d_maps);
    maps_cap = (int)m.size() * 2;
    if (hipMalloc(&d_maps, maps_cap * sizeof(int)) != hipSuccess) return fail(MD_ENOMEM, "maps");
  }
  e = hipMemcpy(d_maps, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "maps");
  e = null_stream_done();
  if (e != hipSuccess) return hip_fail(e, "maps");
  const size_t vneed = 4 * (size_t)std::max(tot[0], 1);
  if (vneed > vpart_cap) {
    if (vpart) (void)hipFree(vpart);
    vpart_cap = 2 * vneed;
    if (hipMalloc(&vpart, vpart_cap * sizeof(double)) != hipSuccess) return fail(MD_ENOMEM, "vpart");
  }
  if (R > steps_cap) {
    if (h_steps) (void)hipHostFree(h_steps);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    h_ctrl = nullptr;
    steps_cap = std::max(2 * R, 8);
    if (hipHostMalloc((void **)&h_steps, steps_cap * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&h_ctrl, steps_cap * sizeof(Ctrl)) != hipSuccess)
      return fail(MD_ENOMEM, "pinned");
  }
  // rebuild queues, one set per position buffer: 2 header words + rep/start
  // ints per entry
  const size_t need_lq = (2 + (size_t)R) + (size_t)R * 3;
  if (2 * need_lq > lq_words) {
    if (d_lq) (void)hipFree(d_lq);
    lq_words = 4 * need_lq;
    if (hipMalloc(&d_lq, lq_words * 8) != hipSuccess) return fail(MD_ENOMEM, "rebuild queue");
  }
  e = hipMemset(d_lq, 0, lq_words * 8);
  if (e == hipSuccess) e = null_stream_done();
  if (e != hipSuccess) return hip_fail(e, "rebuild queue");
  auto grids = [](Plan &P, long nn) {
    P.grid_bin = (int)std::max(1L, std::min<long>(cdiv(nn, mdk::kBlock), 2048));
    P.grid_build = (int)std::max(1L, std::min<long>(cdiv(nn, 8), 2048));
  };
  auto fill = [&](Plan &P, Rep *reps, int nrep, const int *maps, const int *totals) {
    P.reps = reps;
    P.nrep = nrep;
    WMap *w[kMaps] = {&P.part,  &P.chg,   &P.chgG,  &P.fx,    &P.fy,    &P.fz,
                      &P.cell1, &P.cell1w, &P.cell2, &P.cell3, &P.cell3c, &P.brick,
                      &P.plane, &P.partG,  &P.rlst,  &P.cell2s};
    for (int k = 0; k < kMaps; k++) *w[k] = WMap{maps + k * (nrep + 1), nrep, totals[k]};
    P.vpart = vpart;
    P.fp32 = fp32;
    P.real_wg = nrep > 0 && reps_host_wg;
    // the sweep inside the spreading and gather launches pays where it is
    // latency-bound (one system), not in batches (MD_RIDE=0 / =1 forces)
    const char *rd = std::getenv("MD_RIDE");
    P.ride = rd ? rd[0] != '0' : (tune_ride >= 0 ? tune_ride != 0 : !batch);
    P.pspread = pspread;
    const char *rp = std::getenv("MD_RIDE_PARTS");  // forces the plane-spreading ride split
    P.ride_parts = std::max(1, std::min(3, rp ? std::atoi(rp) : tune_rparts));
    const char *rl = std::getenv("MD_RIDE_LEAD");  // forces the riders dispatched first
    P.ride_lead = rl ? std::max(0, std::atoi(rl)) : tune_lead;
    const char *rg2 = std::getenv("MD_RIDE_LEAD_G");
    P.ride_lead_g = rg2 ? std::max(0, std::atoi(rg2)) : tune_lead_g;
    // MD_TWO_STREAMS=1: the real-space sweep beside the SPME kernels on a
    // second stream (the same results; slower in graph replay at the
    // default brush, so off unless asked for)
    const char *two = std::getenv("MD_TWO_STREAMS");
    P.st2 = two && two[0] == '1' ? st2 : nullptr;
    P.ev_fork = ev_fork;
    P.ev_join = ev_join;
  };
  long nn_all = 0;
  for (Sys *s : sys) nn_all += s->p.NN;
  for (int q = 0; q < 2; q++) {
    Plan &B = planp[q];
    unsigned long long *lq = d_lq + (size_t)q * need_lq, *lq_nx = d_lq + (size_t)(q ^ 1) * need_lq;
    Rep *tab = d_reps + (size_t)q * reps_cap;
    fill(B, tab, R, d_maps, tot);
    B.lq = lq;
    B.lq_nx = lq_nx;
    B.lq_cap = R;
    grids(B, nn_all);
    B.lds_x = B.lds_y = B.lds_z = B.lds_sort = B.lds_sort3 = B.lds_plane = 0;
    for (Sys *s : sys) {
      B.lds_x = std::max(B.lds_x, s->lds_x);
      B.lds_y = std::max(B.lds_y, s->lds_y);
      B.lds_z = std::max(B.lds_z, s->lds_z);
      B.lds_plane = std::max(B.lds_plane, s->lds_plane);
      B.lds_sort = std::max(B.lds_sort, s->lds_sort);
      B.lds_sort3 = std::max(B.lds_sort3, s->lds_sort3);
    }
    std::vector<Plan> &one = singlep[q];
    one.assign(R, Plan{});
    for (int r = 0; r < R; r++) {
      fill(one[r], tab + r, 1, d_maps + single_off + r * kMaps * 2, sys[r]->nb);
      one[r].lq = lq + (2 + R) + 3 * (size_t)r;
      one[r].lq_nx = lq_nx + (2 + R) + 3 * (size_t)r;
      one[r].lq_cap = 1;
      grids(one[r], sys[r]->p.NN);
      one[r].lds_x = sys[r]->lds_x;
      one[r].lds_y = sys[r]->lds_y;
      one[r].lds_z = sys[r]->lds_z;
      one[r].lds_plane = sys[r]->lds_plane;
      one[r].lds_sort = sys[r]->lds_sort;
      one[r].lds_sort3 = sys[r]->lds_sort3;
    }
  }
  select(cur);
  return MD_OK;
}

int md_engine::sync() {
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "stream");
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "launch");
  return MD_OK;
}

// Device-side status of one replica from its control block.
int md_engine::check_one(const Sys *s, const Ctrl &c, std::string &msg) const {
  char buf[256];
  if (c.err == MD_EBOND) {
    std::snprintf(buf, sizeof buf, "The bond is break off! particle %d step %d", c.err_info + 1,
                  c.err_step);
    msg = buf;
    return MD_EBOND;
  }
  if (c.err == MD_ECELL) {
    std::snprintf(buf, sizeof buf, "particle %d outside the cell grid (x = Lx/2, z < 0 or z >= Lz)",
                  c.err_info + 1);
    msg = buf;
    return MD_ECELL;
  }
  if (c.err == MD_ENONFINITE) {
    std::snprintf(buf, sizeof buf, "non-finite force on particle %d at step %d", c.err_info + 1,
                  c.err_step);
    msg = buf;
    return MD_ENONFINITE;
  }
  if (c.ovf) {
    std::snprintf(buf, sizeof buf,
                  "capacity exceeded (mask %d): LJ cell %d/%d list %d/%d, real-space cell %d/%d, "
                  "real list %d/%d, spreading tile %d",
                  c.ovf, c.max_cell1, s->h.ccap1, c.max_nb1, s->h.nbcap1, c.max_cell2, s->h.ccap3,
                  c.max_nb2, s->h.nbcap2, s->h.capt);
    msg = buf;
    return MD_EOVERFLOW;
  }
  return MD_OK;
}

int md_engine::check_sys(const Sys *s) {
  int rc = sync();
  if (rc) return rc;
  Ctrl c;
  hipError_t e = hipMemcpy(&c, s->h.ctrl, sizeof c, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "control block");
  std::string msg;
  rc = check_one(s, c, msg);
  return rc ? fail(rc, msg) : MD_OK;
}

int md_engine::check_all() {
  // every control block into pinned memory behind the queued work, one wait
  const bool pinned = h_ctrl && (int)sys.size() <= steps_cap && !dirty;
  if (pinned) {
    for (size_t k = 0; k < sys.size(); k++) {
      hipError_t e = hipMemcpyAsync(&h_ctrl[k], sys[k]->h.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) return hip_fail(e, "control block");
    }
  }
  int rc = sync();
  if (rc) return rc;
  for (size_t k = 0; k < sys.size(); k++) {
    Ctrl c;
    if (pinned) {
      c = h_ctrl[k];
    } else {
      hipError_t e = hipMemcpy(&c, sys[k]->h.ctrl, sizeof c, hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "control block");
    }
    std::string msg;
    rc = check_one(sys[k], c, msg);
    if (rc) return fail(rc, msg);
  }
  return MD_OK;
}

int md_engine::add(const md_params *p, int *slot) {
  if (p->multistep != multistep) return fail(MD_EINVAL, "replicas must share multistep");
  if ((p->precision == MD_FP32) != fp32) return fail(MD_EINVAL, "replicas must share precision");
  Sys *s = new Sys();
  int rc = s->setup(p, fp32);
  if (rc) {
    err = s->err;
    delete s;
    return rc;
  }
  s->h.preheat_end = batch ? std::max(0, p->StepNum0) : 0;
  s->set_par(cur);  // its state is uploaded into the engine's current buffer
  sys.push_back(s);
  *slot = next_slot++;
  slots[*slot] = s;
  dirty = true;
  return MD_OK;
}

int md_engine::launch_compute_force(const Plan &P, Sys *s, int step, int noise, const double *eta,
                                    bool in_step) {
  int mode = mdk::SR_HELD | mdk::SR_FIELD | mdk::SR_ANCHOR;
  if (noise == MD_NOISE_ETA) {
    if (!eta) return fail(MD_EINVAL, "eta is null");
    hipError_t e =
        hipMemcpyAsync(s->d_eta, eta, 3 * sizeof(double) * s->p.NN, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "eta upload");
    mode |= mdk::SR_NOISE_ETA;
  } else if (noise == MD_NOISE_PHILOX) {
    mode |= mdk::SR_NOISE_PHILOX;
  }
  if (in_step) mode |= mdk::SR_KICK | mdk::SR_DRMAX;
  if (s->coulomb && step % multistep == 0) {
    if (fp32) mdk::launch_coulomb<float>(P, st, nullptr);
    else mdk::launch_coulomb<double>(P, st, nullptr);
  }
  if (fp32) mdk::launch_short<float>(P, st, mode, step, s->d_eta, nullptr);
  else mdk::launch_short<double>(P, st, mode, step, s->d_eta, nullptr);
  hipError_t e = hipGetLastError();
  if (noise == MD_NOISE_ETA) {  // eta is host memory of the caller
    hipError_t e2 = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
  }
  return e == hipSuccess ? MD_OK : hip_fail(e, "compute_force launch");
}

// One multistep cycle of steps, captured once and replayed for every
// replica: Coulomb evaluation at the cycle's last step, short range with
// the second kick and the next step's kick and drift (SR_DRIFT), list
// maintenance, gated pre-heat rescale.  A replica enters the cycle at a step
// n with n % multistep == 1, so its Coulomb steps are the multiples of
// multistep, as in Compute_Force (compt_acc.f90:841).  Step numbers come
// from each replica's control block (advanced by the list kernel).  Step k
// of the cycle reads position buffer q ^ ((k - 1) & 1); a run starts with
// one kick-drift (run_cycles) that puts its first step's positions in place.
static void launch_cycle(md_engine *e, int q) {
  const int mode = mdk::SR_HELD | mdk::SR_FIELD | mdk::SR_ANCHOR | mdk::SR_NOISE_PHILOX |
                   mdk::SR_KICK | mdk::SR_DRMAX | mdk::SR_DRIFT | mdk::SR_BIN;
  bool any_coulomb = false, any_preheat = false;
  for (Sys *s : e->sys) {
    any_coulomb |= s->coulomb;
    any_preheat |= s->h.preheat_end > 0;
  }
  const bool any_rl = e->any_real_list();
  for (int k = 1; k <= e->multistep; k++) {
    const Plan &P = e->planp[q ^ ((k - 1) & 1)];
    // the Coulomb step's short-range launch assembles and holds the force
    // the evaluation leaves (SR_COUL) in place of a store kernel
    const bool coul = any_coulomb && k == e->multistep;
    if (coul) {
      if (e->fp32) mdk::launch_coulomb<float>(P, e->st, nullptr, nullptr, false);
      else mdk::launch_coulomb<double>(P, e->st, nullptr, nullptr, false);
    }
    const int mk = coul ? mode | mdk::SR_COUL : mode;
    if (e->fp32) mdk::launch_short<float>(P, e->st, mk, -1, nullptr, nullptr);
    else mdk::launch_short<double>(P, e->st, mk, -1, nullptr, nullptr);
    mdk::launch_build(P, e->st, 1);  // the short-range kernel binned for a rebuild
    if (any_rl) mdk::launch_real_list(P, e->st, false);
    // pre-heat rescale between steps (its step is a multiple of 20, so of
    // multistep when that divides 20; otherwise checked after every step);
    // it redoes the next step's drift with the rescaled velocities
    if (any_preheat && (20 % e->multistep != 0 || k == e->multistep))
      mdk::launch_rescale(P, e->st, 1, 1);
  }
}

int md_engine::build_graph(int q) {
  if (gexec[q]) return MD_OK;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "begin capture");
  launch_cycle(this, q);
  e = hipStreamEndCapture(st, &graph[q]);
  if (e != hipSuccess) return hip_fail(e, "end capture");
  e = hipGraphInstantiate(&gexec[q], graph[q], nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_fail(e, "instantiate");
  e = hipGraphUpload(gexec[q], st);  // device-side set-up before the first replay
  if (e != hipSuccess) return hip_fail(e, "graph upload");
  return MD_OK;
}

// ncycles cycles from the current state: the first step's kick-drift (its
// velocities kept: the step's short-range kernel redoes the kick), then the
// cycles: the first as a replay of its captured graph, which one call
// submits, the others issued on the stream launch by launch while it runs.
// On this runtime a cycle issued on the stream leaves shorter gaps between
// its dependent kernels than its graph replay (the default brush over
// 2,000 steps: 62.7k against 61.2k steps/s; the 256-replica serving
// stream, device-bound, the same either way), but the host, at about 3 us
// per launch, can fall behind the device at the start of a run; the first
// cycle's replay gives it the lead.  Over 20 steps (eight runs each): this
// mix 58.6k-59.3k steps/s, every cycle on the stream 52.5k-60.7k, every
// cycle replayed 55.9k-58.3k.  MD_GRAPH=1 replays every cycle, MD_GRAPH=0
// issues every cycle on the stream (for tracing tools that do not follow
// graph launches).  The last step's positions become the current buffer.
int md_engine::run_cycles(int ncycles) {
  if (ncycles <= 0 || sys.empty()) return MD_OK;
  if (dirty) {
    int rc = rebuild();
    if (rc) return rc;
  }
  const char *gr = std::getenv("MD_GRAPH");
  const bool replay_all = gr && gr[0] == '1', eager = gr && gr[0] == '0';
  const int q0 = cur;
  if (!eager) {
    // both starting buffers: a run of even cycles starts on q0, but it ends
    // with the other buffer current, where the next run starts
    for (int q = 0; q < 2; q++) {
      int rc = build_graph(q);
      if (rc) return rc;
    }
  }
  mdk::launch_kick_drift(planp[q0], st, 0);
  for (int c = 0; c < ncycles; c++) {
    const int q = q0 ^ ((c * multistep) & 1);
    if (eager || (c > 0 && !replay_all)) {
      launch_cycle(this, q);
      continue;
    }
    hipError_t e = hipGraphLaunch(gexec[q], st);
    if (e != hipSuccess) return hip_fail(e, "graph launch");
  }
  select(q0 ^ (((long)ncycles * multistep - 1) & 1));
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MD_OK : hip_fail(e, "cycle launch");
}

// ------------------------------------------------------------------ C-ABI

#define ENGINE_CHECK(e) \
  if (!(e)) return MD_EINVAL
#define FIRST(e)                                                       \
  if ((e)->sys.empty()) return (e)->fail(MD_EINVAL, "engine is empty"); \
  if ((e)->dirty) {                                                    \
    int rc_ = (e)->rebuild();                                          \
    if (rc_) return rc_;                                               \
  }                                                                    \
  Sys *s = (e)->sys[0];                                                \
  const Plan &P = (e)->single[0];                                      \
  (void)P
#define STATE_CHECK(e) \
  if (!s->uploaded) return (e)->fail(MD_EINVAL, "no state uploaded")

static int upload_sys(md_engine *e, Sys *s, const double *pos, const double *vel,
                      const double *acc) {
  const int NN = s->p.NN;
  std::vector<double4> p4(NN);
  std::vector<float4> p32(NN);
  for (int i = 0; i < NN; i++) {
    p4[i] = make_double4(pos[i], pos[NN + i], pos[2 * (size_t)NN + i], pos[3 * (size_t)NN + i]);
    p32[i] = make_float4((float)p4[i].x, (float)p4[i].y, (float)p4[i].z, (float)p4[i].w);
  }
  // no kernel may still use the list buffers that size_lists may replace
  // (a replica not yet uploaded is in no launch: the device keeps running)
  int rc = s->uploaded ? e->sync() : MD_OK;
  if (rc) return rc;
  rc = s->size_lists(pos);
  if (rc) return e->fail(rc, s->err);
  e->dirty = true;
  hipError_t err = hipMemcpy(s->h.pos, p4.data(), NN * sizeof(double4), hipMemcpyHostToDevice);
  if (err == hipSuccess && s->fp32)
    err = hipMemcpy(s->h.pos32, p32.data(), NN * sizeof(float4), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(s->h.vx, vel, 3 * sizeof(double) * NN, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    if (acc) err = hipMemcpy(s->h.ax, acc, 3 * sizeof(double) * NN, hipMemcpyHostToDevice);
    else err = hipMemset(s->h.ax, 0, 3 * sizeof(double) * NN);
  }
  if (err == hipSuccess) err = null_stream_done();
  if (err != hipSuccess) return e->hip_fail(err, "upload");
  s->uploaded = true;
  s->step_dev = -1;
  return MD_OK;
}

static int download_sys(md_engine *e, Sys *s, double *pos, double *vel, double *acc, Ctrl *c) {
  const int NN = s->p.NN;
  int rc = e->sync();
  if (rc) return rc;
  hipError_t err = hipSuccess;
  if (pos) {
    std::vector<double4> p4(NN);
    err = hipMemcpy(p4.data(), s->h.pos, NN * sizeof(double4), hipMemcpyDeviceToHost);
    if (err != hipSuccess) return e->hip_fail(err, "positions");
    for (int i = 0; i < NN; i++) {
      pos[i] = p4[i].x;
      pos[NN + i] = p4[i].y;
      pos[2 * (size_t)NN + i] = p4[i].z;
      pos[3 * (size_t)NN + i] = p4[i].w;
    }
  }
  if (vel) err = hipMemcpy(vel, s->h.vx, 3 * sizeof(double) * NN, hipMemcpyDeviceToHost);
  if (err == hipSuccess && acc) err = hipMemcpy(acc, s->h.ax, 3 * sizeof(double) * NN, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "velocities / forces");
  Ctrl cc;
  err = hipMemcpy(&cc, s->h.ctrl, sizeof cc, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "control block");
  if (c) *c = cc;
  std::string msg;
  rc = e->check_one(s, cc, msg);
  return rc ? e->fail(rc, msg) : MD_OK;
}

extern "C" int md_create(md_engine **out, const md_params *p) {
  if (!out || !p) return MD_EINVAL;
  md_engine *e = new md_engine();
  *out = e;
  int rc = e->init(p->precision, p->multistep);
  if (rc) return rc;
  int slot;
  rc = e->add(p, &slot);
  if (rc) return rc;
  return e->rebuild();
}

extern "C" void md_destroy(md_engine *e) { delete e; }

extern "C" int md_last_error(md_engine *e, char *buf, int32_t len) {
  if (!e || !buf || len <= 0) return MD_EINVAL;
  std::snprintf(buf, (size_t)len, "%s", e->err.c_str());
  return MD_OK;
}

extern "C" int md_upload_state(md_engine *e, const double *pos, const double *vel,
                               const double *acc) {
  ENGINE_CHECK(e && pos && vel);
  FIRST(e);
  int rc = upload_sys(e, s, pos, vel, acc);
  if (rc) return rc;
  return e->rebuild();
}

extern "C" int md_download_state(md_engine *e, double *pos, double *vel, double *acc,
                                 double *dr_max) {
  ENGINE_CHECK(e);
  FIRST(e);
  Ctrl c{};
  int rc = download_sys(e, s, pos, vel, acc, &c);
  if (dr_max) {
    dr_max[0] = c.dr_max1;
    dr_max[1] = c.dr_max2;
  }
  return rc;
}

extern "C" int md_download_coulomb(md_engine *e, double *acc_c) {
  ENGINE_CHECK(e && acc_c);
  FIRST(e);
  int rc = e->sync();
  if (rc) return rc;
  const size_t n = 3 * (size_t)s->p.NN;
  hipError_t err;
  if (s->fp32) {
    std::vector<float> f(n);
    err = hipMemcpy(f.data(), s->h.ccx, n * 4, hipMemcpyDeviceToHost);
    for (size_t k = 0; k < n; k++) acc_c[k] = f[k];
  } else {
    err = hipMemcpy(acc_c, s->h.ccx, n * 8, hipMemcpyDeviceToHost);
  }
  if (err != hipSuccess) return e->hip_fail(err, "held Coulomb force");
  return e->check_all();
}

extern "C" int md_build_lists(md_engine *e) {
  ENGINE_CHECK(e);
  FIRST(e);
  STATE_CHECK(e);
  if (s->h.real_verlet && !s->real_list_wanted) {  // size the real-space list first
    int rc = e->sync();
    if (rc) return rc;
    const int NN = s->p.NN;
    std::vector<double4> p4(NN);
    hipError_t err = hipMemcpy(p4.data(), s->h.pos, NN * sizeof(double4), hipMemcpyDeviceToHost);
    if (err != hipSuccess) return e->hip_fail(err, "positions");
    std::vector<double> pos(4 * (size_t)NN);
    for (int i = 0; i < NN; i++) {
      pos[i] = p4[i].x;
      pos[NN + i] = p4[i].y;
      pos[2 * (size_t)NN + i] = p4[i].z;
      pos[3 * (size_t)NN + i] = p4[i].w;
    }
    s->real_list_wanted = true;
    rc = s->size_lists(pos.data());
    if (rc) return e->fail(rc, s->err);
    e->dirty = true;
    rc = e->rebuild();
    if (rc) return rc;
  }
  const Plan &Q = e->single[0];
  mdk::launch_set_flags(Q, e->st, 0);
  mdk::launch_lists(Q, e->st, 0);
  mdk::launch_real_list(Q, e->st, true);
  mdk::launch_set_flags(Q, e->st, 1);
  return e->check_all();
}

extern "C" int md_download_list(md_engine *e, int32_t which, int32_t *pairs, int64_t cap,
                                int64_t *npair) {
  ENGINE_CHECK(e && npair && (which == 0 || which == 1));
  FIRST(e);
  int rc = e->check_all();
  if (rc) return rc;
  const bool lj = which == 0;
  const int n = lj ? s->p.NN : s->h.Nq;
  const int lcap = lj ? s->h.nbcap1 : s->h.nbcap2;
  if (!lj && !s->h.real_verlet) {
    *npair = 0;
    return MD_OK;
  }
  std::vector<int> cnt(n), nb((size_t)n * lcap);
  hipError_t err = hipMemcpy(cnt.data(), lj ? s->h.nbn1 : s->h.nbn2, n * sizeof(int), hipMemcpyDeviceToHost);
  if (err == hipSuccess)
    err = hipMemcpy(nb.data(), lj ? s->h.nb1 : s->h.nb2, nb.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "lists");
  std::vector<std::pair<int, int>> pr;
  for (int t = 0; t < n; t++) {
    const int i = lj ? t : s->charge[t];
    for (int k = 0; k < std::min(cnt[t], lcap); k++) pr.emplace_back(i + 1, nb[(size_t)k * n + t] + 1);
  }
  std::sort(pr.begin(), pr.end());
  *npair = (int64_t)pr.size();
  if (pairs)
    for (int64_t k = 0; k < (int64_t)pr.size() && k < cap; k++) {
      pairs[k] = pr[k].first;
      pairs[cap + k] = pr[k].second;
    }
  return MD_OK;
}

extern "C" int md_compute_force(md_engine *e, int32_t step, int32_t noise, const double *eta) {
  ENGINE_CHECK(e && step >= 0);
  FIRST(e);
  STATE_CHECK(e);
  return e->launch_compute_force(P, s, step, noise, eta, false);
}

extern "C" int md_new_position(md_engine *e, int32_t step, int32_t noise, const double *eta) {
  ENGINE_CHECK(e && step >= 0);
  FIRST(e);
  STATE_CHECK(e);
  mdk::launch_kick_drift(P, e->st);
  return e->launch_compute_force(P, s, step, noise, eta, true);
}

extern "C" int md_update_verlet_list(md_engine *e) {
  ENGINE_CHECK(e);
  FIRST(e);
  STATE_CHECK(e);
  mdk::launch_lists(P, e->st, 0);
  if (s->h.real_list) mdk::launch_real_list(P, e->st, false);
  mdk::launch_set_flags(P, e->st, 1);
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? MD_OK : e->hip_fail(err, "update_verlet_list");
}

extern "C" int md_rescale_velocity(md_engine *e) {
  ENGINE_CHECK(e);
  FIRST(e);
  STATE_CHECK(e);
  mdk::launch_rescale(P, e->st);
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? MD_OK : e->hip_fail(err, "rescale_velocity");
}

extern "C" int md_run(md_engine *e, int32_t step0, int32_t nsteps, int32_t preheat) {
  ENGINE_CHECK(e && step0 >= 0 && nsteps >= 0);
  FIRST(e);
  STATE_CHECK(e);
  if (e->sys.size() != 1) return e->fail(MD_EINVAL, "md_run drives single-system engines");
  int step = step0;
  const int end = step0 + nsteps;
  auto one = [&](int st) -> int {
    int rc = md_new_position(e, st, MD_NOISE_PHILOX, nullptr);
    if (rc) return rc;
    if (preheat && st % 20 == 0 && st < 300000) {
      rc = md_rescale_velocity(e);
      if (rc) return rc;
    }
    return md_update_verlet_list(e);
  };
  // pre-heat: the replayed cycles rescale through the gated kernels, up to
  // this call's last step; the descriptor goes back to no pre-heat after
  struct Restore {
    md_engine *e;
    Sys *s;
    bool on;
    ~Restore() {
      if (on) {
        s->h.preheat_end = 0;
        e->dirty = true;
      }
    }
  } restore{e, s, preheat != 0};
  if (preheat) {
    s->h.preheat_end = end - 1;
    int rc = e->rebuild();
    if (rc) return rc;
  }
  const int ms = e->multistep;
  while (step < end && (step - 1) % ms != 0) {
    int rc = one(step++);
    if (rc) return rc;
  }
  if (end - step >= ms) {
    // the replayed steps read their numbers from the control block, which
    // the cycles advance: uploaded unless it already holds this step
    if (s->step_dev != step) {
      e->h_steps[0] = step;
      hipError_t err = hipMemcpyAsync(&s->h.ctrl->step, e->h_steps, sizeof(int),
                                      hipMemcpyHostToDevice, e->st);
      if (err != hipSuccess) return e->hip_fail(err, "step upload");
    }
    const int nc = (end - step) / ms;
    s->step_dev = -1;
    int rc = e->run_cycles(nc);
    if (rc) return rc;
    step += nc * ms;
    s->step_dev = step;
  }
  while (step < end) {
    int rc = one(step++);
    if (rc) return rc;
  }
  return e->check_all();
}

extern "C" int md_force_parts(md_engine *e, double *f_short, double *f_coul) {
  ENGINE_CHECK(e);
  FIRST(e);
  STATE_CHECK(e);
  const size_t n = 3 * (size_t)s->p.NN;
  if (e->fp32) mdk::launch_short<float>(P, e->st, 0, 0, nullptr, s->d_out1);
  else mdk::launch_short<double>(P, e->st, 0, 0, nullptr, s->d_out1);
  (void)hipMemsetAsync(s->d_out2, 0, n * sizeof(double), e->st);
  if (s->coulomb) {
    if (e->fp32) mdk::launch_coulomb<float>(P, e->st, s->d_out2);
    else mdk::launch_coulomb<double>(P, e->st, s->d_out2);
  }
  int rc = e->check_all();
  hipError_t err = hipSuccess;
  if (f_short) err = hipMemcpy(f_short, s->d_out1, n * sizeof(double), hipMemcpyDeviceToHost);
  if (err == hipSuccess && f_coul) err = hipMemcpy(f_coul, s->d_out2, n * sizeof(double), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "force parts");
  return rc;
}

// Mean duration of each kernel of a step, from HIP events on the engine's
// stream, over all replicas.  ms[k]: 0 kick_drift (the stepwise form; in
// replayed steps the short-range kernel drifts), 1 short_range (the replayed
// form: second kick and the next step's kick and drift), 2 bin_q, 3 sort_q,
// 4 spread, 5 fft_x_fwd, 6 fft_y_fwd, 7 fft_z_conv, 8 fft_y_inv, 9 fft_x_inv,
// 10 gather, 11 real, 12 bin_lj, 13 (0: the rank pass is part of
// build_lj), 14 build_lj.  Positions,
// velocities and forces are restored; the lists are rebuilt and the held
// Coulomb force recomputed at the current positions.
extern "C" int md_time_kernels(md_engine *e, int32_t iters, double *ms, int32_t n) {
  ENGINE_CHECK(e && ms && iters > 0 && n >= 15);
  FIRST(e);
  STATE_CHECK(e);
  int rc = e->check_all();
  if (rc) return rc;
  const Plan &B = e->plan;
  bool any_coulomb = false;
  for (Sys *x : e->sys) any_coulomb |= x->coulomb;
  hipEvent_t ev[12];
  for (auto &x : ev) (void)hipEventCreate(&x);
  double acc[15] = {0};
  float t;
  auto span = [&](int a, int b) {
    (void)hipEventSynchronize(ev[b]);
    (void)hipEventElapsedTime(&t, ev[a], ev[b]);
    return (double)t;
  };
  // What the timed kick-drift and short-range launches write: both position
  // buffers, velocities, forces, the group displacement maxima, the control
  // block and the rebuild queues; saved once and restored after every
  // iteration of the short-range timing, which runs first, all iterations
  // together, so that no list kernel advances the control block's rebuild
  // generation in between (a restored generation behind the count buffers
  // would make the next binning append to a buffer it did not clear).  The
  // timed launches start from a copy of the control block whose accumulated
  // displacements are -infinity, so that none of them can take a rebuild (a
  // rebuild step bins into the next list generation): they time the
  // cycle's launch on a step without a rebuild, the same instructions.
  const int q0 = e->cur;
  struct Part {
    void *dev;
    size_t n;
  };
  auto parts = [&](Sys *x) {
    const size_t NN = x->p.NN;
    std::vector<Part> v = {{x->posb[0], 32 * NN},  {x->posb[1], 32 * NN},
                           {x->pos32b[0], 16 * NN}, {x->pos32b[1], 16 * NN},
                           {x->h.vx, 24 * NN},      {x->h.ax, 24 * NN},
                           {x->h.dpart, 2 * sizeof(double2) * (size_t)x->h.ngrp},
                           {x->h.ctrl, sizeof(Ctrl)}};
    return v;
  };
  auto rnd = [](size_t n) { return (n + 255) & ~(size_t)255; };
  const size_t R = e->sys.size();
  std::vector<char *> snap(R + 1, nullptr);  // per replica: its parts, then its timing control block
  std::vector<char *> tctrl(R, nullptr);
  auto release = [&]() {
    for (char *x : snap)
      if (x) (void)hipFree(x);
    for (auto &x : ev) (void)hipEventDestroy(x);
  };
  for (size_t k = 0; k <= R; k++) {
    size_t bytes = e->lq_words * 8;
    if (k < R) {
      bytes = 0;
      for (const Part &pt : parts(e->sys[k])) bytes += rnd(pt.n);
      bytes += rnd(sizeof(Ctrl));
    }
    if (hipMalloc((void **)&snap[k], bytes) != hipSuccess) {
      snap[k] = nullptr;
      release();
      return e->fail(MD_ENOMEM, "time_kernels snapshot");
    }
  }
  // the copies' first failure (a failed restore would leave the engine's
  // state changed: the pass then reports it)
  hipError_t copy_err = hipSuccess;
  auto note = [&](hipError_t x) {
    if (copy_err == hipSuccess) copy_err = x;
  };
  auto keep = [&](bool restore, bool timing_ctrl) {
    for (size_t k = 0; k < R; k++) {
      char *h = snap[k];
      for (const Part &pt : parts(e->sys[k])) {
        if (pt.dev) {
          const void *src = restore ? (const void *)h : pt.dev;
          if (restore && timing_ctrl && pt.dev == e->sys[k]->h.ctrl) src = tctrl[k];
          note(hipMemcpyAsync(restore ? pt.dev : (void *)h, src, pt.n, hipMemcpyDeviceToDevice,
                              e->st));
        }
        h += rnd(pt.n);
      }
    }
    char *lq = snap[R];
    if (restore) note(hipMemcpyAsync(e->d_lq, lq, e->lq_words * 8, hipMemcpyDeviceToDevice, e->st));
    else note(hipMemcpyAsync(lq, e->d_lq, e->lq_words * 8, hipMemcpyDeviceToDevice, e->st));
  };
  keep(false, false);
  for (size_t k = 0; k < R; k++) {  // the control blocks the timed launches start from
    size_t off = 0;
    for (const Part &pt : parts(e->sys[k])) off += rnd(pt.n);
    tctrl[k] = snap[k] + off;
    Ctrl c;
    hipError_t err = hipMemcpyAsync(&c, e->sys[k]->h.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, e->st);
    if (err == hipSuccess) err = hipStreamSynchronize(e->st);
    c.dr_max1 = c.dr_max2 = -HUGE_VAL;
    c.pend = 0;
    if (err == hipSuccess) err = hipMemcpy(tctrl[k], &c, sizeof(Ctrl), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
      release();
      return e->hip_fail(err, "time_kernels control block");
    }
  }
  // the launch of a replayed step (launch_cycle): list bookkeeping, the
  // second kick, the next step's kick and drift into the other buffer
  const int mode = mdk::SR_HELD | mdk::SR_FIELD | mdk::SR_ANCHOR | mdk::SR_NOISE_PHILOX |
                   mdk::SR_KICK | mdk::SR_DRMAX | mdk::SR_DRIFT | mdk::SR_BIN;
  for (int it = 0; it < iters; it++) {
    // timed over back-to-back launches, so that the per-launch mean
    // excludes the event overhead; each launch reads the buffer the one
    // before wrote, as in the cycle; the kick-drift that starts a run as
    // the difference between [kick, shorts] and [shorts]
    constexpr int kRepeat = 8;
    auto shorts = [&](int k0) {
      for (int k = 0; k < kRepeat; k++) {
        const Plan &Pk = e->planp[q0 ^ ((k0 + k) & 1)];
        if (e->fp32) mdk::launch_short<float>(Pk, e->st, mode, -1, nullptr, nullptr);
        else mdk::launch_short<double>(Pk, e->st, mode, -1, nullptr, nullptr);
      }
    };
    keep(true, true);
    (void)hipEventRecord(ev[0], e->st);
    mdk::launch_kick_drift(e->planp[q0], e->st, 0);
    shorts(0);
    (void)hipEventRecord(ev[1], e->st);
    shorts(kRepeat);
    (void)hipEventRecord(ev[2], e->st);
    const double both = span(0, 1), sh = span(1, 2);
    acc[0] += std::max(0.0, both - sh);
    acc[1] += sh / kRepeat;
  }
  keep(true, false);  // the state as it was; the Coulomb and list kernels move no particle
  for (int it = 0; it < iters; it++) {
    if (any_coulomb) {
      if (e->fp32) mdk::launch_coulomb<float>(B, e->st, nullptr, ev);
      else mdk::launch_coulomb<double>(B, e->st, nullptr, ev);
      for (int k = 0; k < 10; k++) acc[2 + k] += span(k, k + 1);
    }
    mdk::launch_set_flags(B, e->st, 0);
    mdk::launch_lists(B, e->st, 0, ev);
    for (int k = 0; k < 3; k++) acc[12 + k] += span(k, k + 1);
  }
  mdk::launch_set_flags(B, e->st, 1);
  rc = e->check_all();
  for (char *x : snap) (void)hipFree(x);
  for (auto &x : ev) (void)hipEventDestroy(x);
  for (int k = 0; k < 15; k++) ms[k] = acc[k] / iters;
  if (rc == MD_OK && copy_err != hipSuccess) return e->hip_fail(copy_err, "time_kernels state copies");
  return rc;
}

// Chooses the launch forms (real-space sweep form and whether it shares the
// spreading and gather launches, z-pass lines per workgroup, fused FFT
// planes) by timing one Coulomb evaluation of the
// engine's replicas under each: the forms give the same results, so only
// the time differs.  Knobs are tuned one after the other, `iters`
// evaluations each.
extern "C" int md_autotune(md_engine *e, int32_t iters, int32_t *choice) {
  ENGINE_CHECK(e && iters > 0);
  FIRST(e);
  STATE_CHECK(e);
  bool any_coulomb = false;
  for (Sys *x : e->sys) any_coulomb |= x->coulomb;
  if (!any_coulomb) return MD_OK;
  int rc = e->check_all();
  if (rc) return rc;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  auto time_coulomb = [&](double &ms) -> int {
    int r2 = e->rebuild();
    if (r2) return r2;
    const Plan &B = e->plan;
    auto one = [&]() {
      if (e->fp32) mdk::launch_coulomb<float>(B, e->st, nullptr);
      else mdk::launch_coulomb<double>(B, e->st, nullptr);
    };
    one();  // warm
    // the best of three batches of `iters` evaluations: a batch that a
    // host or device hiccup slowed does not decide the choice
    ms = 1e30;
    for (int rep = 0; rep < 3; rep++) {
      (void)hipEventRecord(a, e->st);
      for (int k = 0; k < iters; k++) one();
      (void)hipEventRecord(b, e->st);
      (void)hipEventSynchronize(b);
      float t = 0;
      (void)hipEventElapsedTime(&t, a, b);
      ms = std::min(ms, (double)t / iters);
    }
    return MD_OK;
  };
  struct Knob {
    int *field;
    std::vector<int> options;
  };
  Knob knobs[] = {{&e->tune_fuse, {0, 1}}, {&e->tune_pz, {2, 4, 8, 16}}, {&e->tune_real, {0, 1}},
                  {&e->tune_ride, {0, 1}}, {&e->tune_rgrid, {1024, 2048}}, {&e->tune_pspread, {0, 1}},
                  {&e->tune_rparts, {1, 2, 3}}, {&e->tune_lead, {0, 256, 512, 768}},
                  {&e->tune_lead_g, {0, 512, 1024, 65536}}};
  for (Knob &k : knobs) {
    int best = *k.field;
    double best_ms = 1e30;
    for (int opt : k.options) {
      *k.field = opt;
      double ms = 0;
      rc = time_coulomb(ms);
      if (rc) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = opt;
      }
    }
    *k.field = best;
    if (rc) break;
  }
  e->dirty = true;
  if (rc == MD_OK) rc = e->check_all();
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (choice) {
    choice[0] = e->tune_fuse;
    choice[1] = e->tune_pz;
    choice[2] = e->tune_real;
    choice[3] = e->tune_ride;
    choice[4] = e->tune_rgrid;
    choice[5] = e->tune_pspread;
    choice[6] = e->tune_rparts;
    choice[7] = e->tune_lead;
    choice[8] = e->tune_lead_g;
  }
  return rc;
}

// Chooses the engine geometry of a system by running it: the spreading
// tile edges (the spread sums are exact fixed point, so results do not
// change) and, with flags & MD_TUNE_LISTS, the LJ Verlet skin and the
// real-space sweep's cell scale (the same pairs within every cut, but the
// lists and sweep cells order the sums differently: forces change by
// rounding).  Each candidate is a fresh engine started from (pos, vel):
// lists, step-0 forces, warm-up cycles (graph capture and first replays),
// then `cycles` multistep cycles timed on the host around md_run, which ends
// with a stream synchronise.  The knobs are tuned one after the other, each
// keeping the best so far; a candidate replaces the incumbent only when it
// is faster by kMargin (results-changing knobs by kMarginLists), so timing
// noise does not move the reference's choice.
extern "C" int md_autotune_geometry(const md_params *p, const double *pos, const double *vel,
                                    int32_t cycles, int32_t flags, md_params *best,
                                    double *steps_per_s) {
  if (!p || !pos || !vel || !best || cycles <= 0) return MD_EINVAL;
  constexpr double kMargin = 1.01, kMarginLists = 1.03;
  constexpr int kWarmCycles = 10;
  auto rate = [&](const md_params &cand, double &sps) -> int {
    md_params d = cand;
    int rc = md_derive_params(&d);
    if (rc) return rc;
    md_engine *e = nullptr;
    rc = md_create(&e, &d);
    if (!rc) rc = md_upload_state(e, pos, vel, nullptr);
    if (!rc) rc = md_build_lists(e);
    if (!rc) rc = md_compute_force(e, 0, MD_NOISE_PHILOX, nullptr);
    const int ms = std::max(1, d.multistep);
    if (!rc) rc = md_run(e, 1, kWarmCycles * ms, 0);
    if (!rc) {
      const auto t0 = std::chrono::steady_clock::now();
      rc = md_run(e, 1 + kWarmCycles * ms, cycles * ms, 0);
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      sps = dt > 0 ? cycles * ms / dt : 0;
    }
    md_destroy(e);
    return rc;
  };
  md_params cur = *p;
  double cur_sps = 0;
  int rc = rate(cur, cur_sps);
  if (rc) return rc;
  auto consider = [&](const md_params &cand, double margin) {
    double sps = 0;
    if (rate(cand, sps) == MD_OK && sps > cur_sps * margin) {
      cur = cand;
      cur_sps = sps;
    }
  };
  // tile edges: 16 or 8 points along x and y together, then along z
  const int start_tile[3] = {p->ovr_tile[0], p->ovr_tile[1], p->ovr_tile[2]};
  for (int e : {16, 8}) {
    md_params c = cur;
    c.ovr_tile[0] = c.ovr_tile[1] = e;
    if (e != (start_tile[0] ? start_tile[0] : 16)) consider(c, kMargin);
  }
  for (int e : {16, 8}) {
    md_params c = cur;
    c.ovr_tile[2] = e;
    if (e != (start_tile[2] ? start_tile[2] : 16)) consider(c, kMargin);
  }
  if (flags & MD_TUNE_LISTS) {
    for (double skin : {0.75, 1.0}) {  // the reference's skin is 0.5 (rvl = rcl_s + 1)
      md_params c = cur;
      c.ovr_skin = skin;
      consider(c, kMarginLists);
    }
    for (double scale : {1.25, 1.5}) {
      md_params c = cur;
      c.ovr_cell_scale = scale;
      consider(c, kMarginLists);
    }
  }
  rc = md_derive_params(&cur);
  if (rc) return rc;
  *best = cur;
  if (steps_per_s) *steps_per_s = cur_sps;
  return MD_OK;
}

// error_analysis (compt_acc.f90:729-812) for one replica.  The production
// estimates f_r, f_k (compt_acc.f90:746-751: exponents (1./3) in default
// real, (Kmax1*Kmax2*Kmax3)**(1./3) evaluated in default real, Kmax the
// mesh); the reference force: the rule at tol = 5, tau_rf = 10 without
// overrides (:756-759), k-ranges ceiling(Kmax/3.) (:760-762),
// Standard_Ewald + real_space over every pair within r_c (:768-770), on the
// device in FP64; against the replica's own Coulomb force (SPME + real
// space + slab, in the engine's precision); rmsf over every particle (the
// reference sums rows 1..Nq, :797-801).
int md_engine::error_analysis(Sys *s, const Plan &P, double *out, double *acc_se) {
  if (!out) return fail(MD_EINVAL, "out is null");
  if (!s->coulomb) return fail(MD_EINVAL, "error_analysis needs a charged system");
  int rc = check_sys(s);
  if (rc) return rc;
  const md_params &p = s->p;
  const int NN = p.NN, Nq = s->h.Nq;
  // production estimates
  std::vector<double4> hp(NN);
  hipError_t err = hipMemcpy(hp.data(), s->h.pos, NN * sizeof(double4), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return hip_fail(err, "positions");
  double qtot = 0;
  for (int m = 0; m < Nq; m++) qtot += hp[s->charge[m]].w * hp[s->charge[m]].w;
  const double kPi = 3.141592653589793, third = (double)(1.0f / 3.0f);
  const double ex = std::exp(-p.tol * p.tol);
  const double f_r = 2 * std::fabs(p.qq) * std::sqrt(qtot / p.rcc / p.Lx / p.Ly / p.Lz) * ex;
  const float kprod = std::pow((float)(p.ng[0] * p.ng[1] * p.ng[2]), 1.0f / 3.0f);
  const double f_k = std::fabs(p.qq) * p.alpha / std::pow(p.Lx * p.Ly * p.Lz, third) / kPi *
                     std::sqrt(8 * qtot / (double)kprod) * ex;
  // parameters of the reference force
  double alpha, rcc, rvc;
  int k[3], rv;
  mdk::ewald_rule(&p, 5.0, 10.0, false, &alpha, &rcc, &rvc, k, &rv);
  int K[3];
  for (int d = 0; d < 3; d++) K[d] = (int)std::ceil((float)(int)std::ceil((float)k[d] * 3.0f) / 3.0f);
  // k-vectors in the reference's order (fourier_function, compt_acc.f90:650-668)
  const double Ze = p.Zfac, c1 = 2 * kPi / p.Lx, c2 = 2 * kPi / p.Ly, c3 = 2 * kPi / p.Lz / Ze;
  std::vector<double4> kv;
  for (int r = 0; r <= K[2]; r++) {
    const double factor = r == 0 ? 1 : 2;
    for (int pp = -K[0]; pp <= K[0]; pp++)
      for (int q = -K[1]; q <= K[1]; q++) {
        const float a = 1.0f * pp / K[0], b = 1.0f * q / K[1], c = 1.0f * r / K[2];
        const float kcut = a * a + b * b + c * c;
        if (kcut > 1 || kcut == 0) continue;
        const double kx = pp * c1, ky = q * c2, kz = r * c3;
        const double ksqr = kx * kx + ky * ky + kz * kz;
        const double ek = factor * 4 * kPi / (p.Lx * p.Ly * p.Lz * Ze) *
                          std::exp(-ksqr / 4 / (alpha * alpha)) / ksqr * p.lb / p.Beta;
        kv.push_back(make_double4(kx, ky, kz, ek));
      }
  }
  mdk::EwaldArgs a;
  a.Nq = Nq;
  a.NN = NN;
  a.nk = (int)kv.size();
  const int bq = cdiv(Nq, 256), bk = cdiv(a.nk, 256);
  a.ms = std::max(1, std::min(16, cdiv(1024, bk)));
  a.ks = std::max(1, std::min(bk, cdiv(1024, bq)));
  a.rs = std::max(1, std::min(bq, cdiv(1024, bq)));
  a.charge = s->h.charge;
  a.pos = s->h.pos;
  a.Lx = p.Lx;
  a.Ly = p.Ly;
  a.hLx = p.Lx / 2;
  a.hLy = p.Ly / 2;
  a.alpha = alpha;
  a.alpha2 = alpha * alpha;
  a.c2 = 2 * alpha / std::sqrt(kPi);
  a.rcc2 = rcc * rcc;
  a.c1 = p.lb / p.Beta;
  a.slab = 4 * kPi / p.Lx / p.Ly / p.Lz / Ze;
  const size_t nq = std::max(Nq, 1), nk = std::max(a.nk, 1);
  const size_t bytes = 32 * nq + 32 * nk + 16 * nk * a.ms + 16 * nk + 24 * nq * (a.ks + a.rs) + 256;
  char *buf = nullptr;
  if (hipMalloc((void **)&buf, bytes) != hipSuccess) return fail(MD_ENOMEM, "error_analysis buffers");
  char *cur_b = buf;
  auto carve = [&](size_t n) {
    char *x = cur_b;
    cur_b += (n + 255) / 256 * 256;
    return x;
  };
  a.cq = (double4 *)carve(32 * nq);
  double4 *dkv = (double4 *)carve(32 * nk);
  a.kv = dkv;
  a.sp = (double2 *)carve(16 * nk * a.ms);
  a.sk = (double2 *)carve(16 * nk);
  a.rp = (double *)carve(24 * nq * a.ks);
  a.qp = (double *)carve(24 * nq * a.rs);
  a.mz = (double *)carve(8);
  a.out = s->d_out1;
  const size_t n3 = 3 * (size_t)NN;
  err = hipMemcpyAsync(dkv, kv.data(), kv.size() * sizeof(double4), hipMemcpyHostToDevice, st);
  if (err == hipSuccess) err = hipMemsetAsync(s->d_out1, 0, n3 * sizeof(double), st);
  if (err == hipSuccess) err = hipMemsetAsync(s->d_out2, 0, n3 * sizeof(double), st);
  if (err == hipSuccess) {
    if (a.nk > 0) mdk::launch_ewald(a, st);
    if (fp32) mdk::launch_coulomb<float>(P, st, s->d_out2);
    else mdk::launch_coulomb<double>(P, st, s->d_out2);
    err = hipStreamSynchronize(st);
  }
  if (err == hipSuccess) err = hipGetLastError();
  std::vector<double> fse(n3), fc(n3);
  if (err == hipSuccess) err = hipMemcpy(fse.data(), s->d_out1, n3 * sizeof(double), hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(fc.data(), s->d_out2, n3 * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (err != hipSuccess) return hip_fail(err, "error_analysis");
  double s1 = 0, s2 = 0;
  for (size_t j = 0; j < n3; j++) {
    s1 += (fse[j] - fc[j]) * (fse[j] - fc[j]);
    s2 += fse[j] * fse[j];
  }
  out[0] = s2 > 0 ? std::sqrt(s1 / s2) : 0;
  out[1] = f_r;
  out[2] = f_k;
  out[3] = alpha;
  out[4] = rcc;
  out[5] = K[0];
  out[6] = K[1];
  out[7] = K[2];
  if (acc_se) std::memcpy(acc_se, fse.data(), n3 * sizeof(double));
  return check_sys(s);
}

// error_analysis without any allocation (for callers that cannot allocate:
// graph capture, tight serving loops): the same reference force, with its
// buffers carved from the library's per-GPU device scratch
// (mdk::ewald_scratch, kEwaldScratchBytes), the k-vectors as the whole box
// with zero weight outside the ellipsoid (the same sums up to their
// grouping), fill kernels in place of memsets, and the rmsf sums and sum q^2
// on the device.  Sizes beyond the scratch are refused.  Calls that overlap
// on one GPU share its scratch and race; every buffer is written before it
// is read, so consecutive calls need no clearing.
int md_engine::error_analysis_noalloc(Sys *s, const Plan &P, double *out, double *acc_se) {
  if (!out) return fail(MD_EINVAL, "out is null");
  if (!s->coulomb) return fail(MD_EINVAL, "error_analysis needs a charged system");
  char *scr = (char *)mdk::ewald_scratch();
  if (!scr) return fail(MD_EHIP, "device scratch address");
  int rc = check_sys(s);
  if (rc) return rc;
  const md_params &p = s->p;
  const int NN = p.NN, Nq = s->h.Nq;
  double alpha, rcc, rvc;
  int k[3], rv;
  mdk::ewald_rule(&p, 5.0, 10.0, false, &alpha, &rcc, &rvc, k, &rv);
  int K[3];
  for (int d = 0; d < 3; d++) K[d] = (int)std::ceil((float)(int)std::ceil((float)k[d] * 3.0f) / 3.0f);
  const double Ze = p.Zfac, c1 = 2 * kPi / p.Lx, c2 = 2 * kPi / p.Ly, c3 = 2 * kPi / p.Lz / Ze;
  const double pref = 4 * kPi / (p.Lx * p.Ly * p.Lz * Ze) * p.lb / p.Beta;
  mdk::EwaldArgs a;
  a.Nq = Nq;
  a.NN = NN;
  const long nbox = (long)(K[2] + 1) * (2 * K[0] + 1) * (2 * K[1] + 1);
  if (nbox > (1L << 28)) return fail(MD_EINVAL, "error_analysis k-space too large for the device scratch");
  a.nk = (int)nbox;
  const int bq = cdiv(Nq, 256), bk = cdiv(a.nk, 256);
  a.ms = std::max(1, std::min(16, cdiv(1024, bk)));
  a.ks = std::max(1, std::min(bk, cdiv(1024, bq)));
  a.rs = std::max(1, std::min(bq, cdiv(1024, bq)));
  a.charge = s->h.charge;
  a.pos = s->h.pos;
  a.Lx = p.Lx;
  a.Ly = p.Ly;
  a.hLx = p.Lx / 2;
  a.hLy = p.Ly / 2;
  a.alpha = alpha;
  a.alpha2 = alpha * alpha;
  a.c2 = 2 * alpha / std::sqrt(kPi);
  a.rcc2 = rcc * rcc;
  a.c1 = p.lb / p.Beta;
  a.slab = 4 * kPi / p.Lx / p.Ly / p.Lz / Ze;
  const size_t nq = std::max(Nq, 1), nk = std::max(a.nk, 1);
  size_t used = 0;
  auto carve = [&](size_t n) {
    char *x = scr + used;
    used += (n + 255) / 256 * 256;
    return x;
  };
  a.cq = (double4 *)carve(32 * nq);
  double4 *kvd = (double4 *)carve(32 * nk);
  a.kv = kvd;
  a.sp = (double2 *)carve(16 * nk * a.ms);
  a.sk = (double2 *)carve(16 * nk);
  a.rp = (double *)carve(24 * nq * a.ks);
  a.qp = (double *)carve(24 * nq * a.rs);
  a.mz = (double *)carve(8);
  double *stats = (double *)carve(3 * 8);
  if (used > mdk::kEwaldScratchBytes)
    return fail(MD_EINVAL, "error_analysis needs " + std::to_string(used) + " bytes of device scratch, more than " +
                               std::to_string(mdk::kEwaldScratchBytes));
  a.out = s->d_out1;
  const size_t n3 = 3 * (size_t)NN;
  mdk::launch_ewald_box(a.nk, kvd, st, K[0], K[1], K[2], c1, c2, c3, pref, alpha * alpha);
  mdk::launch_zero(s->d_out1, n3, st);
  mdk::launch_zero(s->d_out2, n3, st);
  mdk::launch_ewald(a, st);
  if (fp32) mdk::launch_coulomb<float>(P, st, s->d_out2);
  else mdk::launch_coulomb<double>(P, st, s->d_out2);
  mdk::launch_ewald_stats(s->d_out1, s->d_out2, n3, a, stats, st);
  double h[3];
  hipError_t err = hipMemcpyAsync(h, stats, sizeof h, hipMemcpyDeviceToHost, st);
  if (err == hipSuccess && acc_se) err = hipMemcpyAsync(acc_se, s->d_out1, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err == hipSuccess) err = hipGetLastError();
  if (err != hipSuccess) return hip_fail(err, "error_analysis");
  // production estimates (compt_acc.f90:746-751), as error_analysis computes them
  const double third = (double)(1.0f / 3.0f);
  const double ex = std::exp(-p.tol * p.tol);
  const double f_r = 2 * std::fabs(p.qq) * std::sqrt(h[2] / p.rcc / p.Lx / p.Ly / p.Lz) * ex;
  const float kprod = std::pow((float)(p.ng[0] * p.ng[1] * p.ng[2]), 1.0f / 3.0f);
  const double f_k = std::fabs(p.qq) * p.alpha / std::pow(p.Lx * p.Ly * p.Lz, third) / kPi *
                     std::sqrt(8 * h[2] / (double)kprod) * ex;
  out[0] = h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0;
  out[1] = f_r;
  out[2] = f_k;
  out[3] = alpha;
  out[4] = rcc;
  out[5] = K[0];
  out[6] = K[1];
  out[7] = K[2];
  return check_sys(s);
}

extern "C" int md_error_analysis(md_engine *e, double *out, double *acc_se) {
  ENGINE_CHECK(e && out);
  FIRST(e);
  STATE_CHECK(e);
  return e->error_analysis(s, P, out, acc_se);
}

// ------------------------------------------------------------- batches

extern "C" int md_batch_create(md_engine **out, int32_t precision, int32_t multistep) {
  if (!out) return MD_EINVAL;
  md_engine *e = new md_engine();
  *out = e;
  e->batch = true;
  return e->init(precision, multistep);
}

extern "C" int md_batch_add(md_engine *e, const md_params *p, const double *pos, const double *vel,
                            int32_t *slot) {
  ENGINE_CHECK(e && p && pos && vel && slot);
  int sl;
  int rc = e->add(p, &sl);
  if (rc) return rc;
  Sys *s = e->slots[sl];
  rc = upload_sys(e, s, pos, vel, nullptr);
  if (rc == MD_OK) rc = e->rebuild();
  if (rc == MD_OK) {
    const Plan &P = e->single[e->index_of(s)];
    mdk::launch_set_flags(P, e->st, 0);
    mdk::launch_lists(P, e->st, 0);
    if (s->h.real_list) mdk::launch_real_list(P, e->st, false);
    mdk::launch_set_flags(P, e->st, 1);
    rc = e->launch_compute_force(P, s, 0, MD_NOISE_PHILOX, nullptr, false);
  }
  if (rc == MD_OK) {
    int one = 1;  // the first replayed step of this replica
    hipError_t err = hipMemcpyAsync(&s->h.ctrl->step, &one, sizeof one, hipMemcpyHostToDevice, e->st);
    if (err == hipSuccess) err = hipStreamSynchronize(e->st);
    Ctrl c;
    if (err == hipSuccess) err = hipMemcpy(&c, s->h.ctrl, sizeof c, hipMemcpyDeviceToHost);
    if (err != hipSuccess) {
      rc = e->hip_fail(err, "control block");
    } else {  // a request whose initial state is invalid is refused here
      std::string msg;
      rc = e->check_one(s, c, msg);
      if (rc) e->fail(rc, msg);
    }
  }
  if (rc) {
    (void)hipStreamSynchronize(e->st);
    e->slots.erase(sl);
    e->sys.erase(std::find(e->sys.begin(), e->sys.end(), s));
    delete s;
    e->dirty = true;
    return rc;
  }
  *slot = sl;
  return MD_OK;
}

extern "C" int md_batch_remove(md_engine *e, int32_t slot) {
  ENGINE_CHECK(e);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  int rc = e->sync();
  if (rc) return rc;
  Sys *s = it->second;
  e->sys.erase(std::find(e->sys.begin(), e->sys.end(), s));
  e->slots.erase(it);
  delete s;
  e->dirty = true;
  return MD_OK;
}

extern "C" int md_batch_run(md_engine *e, int32_t ncycles) {
  ENGINE_CHECK(e && ncycles >= 0);
  int rc = e->run_cycles(ncycles);
  if (rc) return rc;
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? MD_OK : e->hip_fail(err, "batch run");
}

extern "C" int md_batch_status(md_engine *e, int32_t slot, int32_t *step, char *buf, int32_t len) {
  ENGINE_CHECK(e);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  int rc = e->sync();
  if (rc) return rc;
  Ctrl c;
  hipError_t err = hipMemcpy(&c, it->second->h.ctrl, sizeof c, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "control block");
  if (step) *step = c.step;
  std::string msg;
  rc = e->check_one(it->second, c, msg);
  if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", msg.c_str());
  return rc;
}

extern "C" int md_batch_download(md_engine *e, int32_t slot, double *pos, double *vel, double *acc) {
  ENGINE_CHECK(e);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  return download_sys(e, it->second, pos, vel, acc, nullptr);
}

extern "C" int md_replica_bytes(const md_params *p, const double *pos, int64_t *bytes) {
  if (!p || !pos || !bytes) return MD_EINVAL;
  Sys s;
  s.dry = true;  // every allocation counted, none made; no device call
  int rc = s.setup(p, p->precision == MD_FP32);
  if (rc == MD_OK) rc = s.size_lists(pos);
  *bytes = rc == MD_OK ? (int64_t)s.bytes : 0;
  return rc;
}

extern "C" int md_batch_info(md_engine *e, int32_t slot, int64_t *bytes, int32_t *nrep) {
  ENGINE_CHECK(e);
  if (nrep) *nrep = (int32_t)e->sys.size();
  if (bytes) {
    if (slot < 0) {
      int64_t b = 0;
      for (Sys *s : e->sys) b += (int64_t)s->bytes;
      *bytes = b;
    } else {
      auto it = e->slots.find(slot);
      if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
      *bytes = (int64_t)it->second->bytes;
    }
  }
  return MD_OK;
}

extern "C" int md_batch_observe(md_engine *e, int32_t slot, double *height, double *profile,
                                int32_t nbins) {
  ENGINE_CHECK(e && nbins == mdk::kHistBins);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  Sys *s = it->second;
  if (e->dirty) {
    int rc = e->rebuild();
    if (rc) return rc;
  }
  const int idx = e->index_of(s);
  mdk::launch_observe(e->single[idx], e->st, s->d_hist, s->d_hsum);
  int rc = e->sync();
  if (rc) return rc;
  const int nb = cdiv(s->p.NN, mdk::kBlock);
  std::vector<double> hs(nb);
  std::vector<int> hist(mdk::kHistBins);
  hipError_t err = hipMemcpy(hs.data(), s->d_hsum, nb * sizeof(double), hipMemcpyDeviceToHost);
  if (err == hipSuccess)
    err = hipMemcpy(hist.data(), s->d_hist, hist.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "observables");
  double sum = 0;
  for (double v : hs) sum += v;
  if (height) *height = sum / s->p.Npe;
  if (profile)
    for (int k = 0; k < mdk::kHistBins; k++) profile[k] = hist[k];
  return MD_OK;
}

extern "C" int md_batch_error_analysis_noalloc(md_engine *e, int32_t slot, double *out,
                                               double *acc_se) {
  ENGINE_CHECK(e && out);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  // a membership change re-uploads descriptors (an allocation): run a cycle first
  if (e->dirty) return e->fail(MD_EINVAL, "replicas changed since the last run: run a cycle first");
  Sys *s = it->second;
  return e->error_analysis_noalloc(s, e->single[e->index_of(s)], out, acc_se);
}

extern "C" int md_batch_error_analysis(md_engine *e, int32_t slot, double *out, double *acc_se) {
  ENGINE_CHECK(e && out);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  if (e->dirty) {
    int rc = e->rebuild();
    if (rc) return rc;
  }
  Sys *s = it->second;
  return e->error_analysis(s, e->single[e->index_of(s)], out, acc_se);
}

// The height analysis of write_height (in_output.f90:1017-1079,1342-1344):
// out[8] = h_avg, hl_max, hl_end, hl_avg, hs_max, hs_end, hs_branch,
// hs_avg (the columns after the step of height.txt), from per-molecule
// sums on the device added up here in molecule order.  A molecule kind
// that the system lacks reads 0 (the reference divides by zero there).
static int heights(md_engine *e, Sys *s, const Plan &P, double *out) {
  if (!out) return e->fail(MD_EINVAL, "out is null");
  const md_params &p = s->p;
  const int nm = p.Nga + p.Ngl;
  mdk::launch_observe(P, e->st, s->d_hist, s->d_hsum);
  mdk::launch_heights(P, e->st, s->d_mol);
  int rc = e->check_sys(s);
  if (rc) return rc;
  const int nb = cdiv(p.NN, mdk::kBlock);
  std::vector<double> hs(nb), mol(4 * (size_t)std::max(nm, 1));
  hipError_t err = hipMemcpy(hs.data(), s->d_hsum, nb * sizeof(double), hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(mol.data(), s->d_mol, mol.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "heights");
  double sum = 0;
  for (double v : hs) sum += v;
  double hl_avg = 0, hl_max = 0, hl_end = 0, hs_avg = 0, hs_max = 0, hs_branch = 0, hs_end = 0;
  for (int k = 0; k < p.Ngl; k++) {
    hl_avg += mol[4 * k];
    hl_max += mol[4 * k + 1];
    hl_end += mol[4 * k + 2];
  }
  for (int k = p.Ngl; k < nm; k++) {
    hs_avg += mol[4 * k];
    hs_max += mol[4 * k + 1];
    hs_branch += mol[4 * k + 2];
    hs_end += mol[4 * k + 3];
  }
  out[0] = p.Npe > 0 ? sum / p.Npe : 0;
  out[1] = p.Ngl > 0 ? hl_max / p.Ngl : 0;
  out[2] = p.Ngl > 0 ? hl_end / p.Ngl : 0;
  out[3] = p.Ngl > 0 && p.Nml > 0 ? hl_avg / p.Ngl / p.Nml : 0;
  out[4] = p.Nga > 0 ? hs_max / p.Nga : 0;
  out[5] = p.Nga > 0 && p.arm > 1 ? hs_end / (p.Nga * (p.arm - 1)) : 0;
  out[6] = p.Nga > 0 ? hs_branch / p.Nga : 0;
  out[7] = p.Nga > 0 ? hs_avg / (p.Nga * (p.arm * p.Nma + 1)) : 0;
  return MD_OK;
}

// Energies and pair virials of one replica at its current state (the
// Coulomb evaluation is rerun at these positions; the held Coulomb force and
// the trajectory are unchanged).  out[16]: WCA, wall, FENE, real space,
// reciprocal, slab, Ewald self term, Coulomb (real + reciprocal + slab +
// self), field, potential total, kinetic sum |v|^2/2, the reference's
// kenetic_energy (kinetic / (NN - N_anchor), in_output.f90:1307-1311), the
// pair virials sum r.F of WCA, FENE and real space, 0.
static int energies(md_engine *e, Sys *s, const Plan &P, double *out) {
  if (!out) return e->fail(MD_EINVAL, "out is null");
  int rc = e->check_sys(s);
  if (rc) return rc;
  if (s->coulomb) {
    if (s->fp32) mdk::launch_coulomb<float>(P, e->st, s->d_out2);
    else mdk::launch_coulomb<double>(P, e->st, s->d_out2);
  }
  const int *nb = s->en_blocks;
  if (s->fp32) mdk::launch_energies<float>(P, e->st, s->d_epart, nb[0], nb[1], nb[2]);
  else mdk::launch_energies<double>(P, e->st, s->d_epart, nb[0], nb[1], nb[2]);
  rc = e->check_sys(s);
  if (rc) return rc;
  const size_t n = (size_t)nb[0] * mdk::kEnShortTerms + 2 * (size_t)nb[1] + nb[2];
  std::vector<double> h(n);
  hipError_t err = hipMemcpy(h.data(), s->d_epart, n * sizeof(double), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->hip_fail(err, "energies");
  double t[mdk::kEnShortTerms] = {0}, re = 0, rw = 0, qu = 0;
  for (int b = 0; b < nb[0]; b++)
    for (int k = 0; k < mdk::kEnShortTerms; k++) t[k] += h[(size_t)b * mdk::kEnShortTerms + k];
  const double *h2 = h.data() + (size_t)nb[0] * mdk::kEnShortTerms;
  for (int b = 0; b < nb[1]; b++) {
    re += h2[2 * b];
    rw += h2[2 * b + 1];
  }
  for (int b = 0; b < nb[2]; b++) qu += h2[2 * (size_t)nb[1] + b];
  const md_params &p = s->p;
  const double c1 = p.lb / p.Beta;
  std::memset(out, 0, 16 * sizeof(double));
  out[0] = t[0];
  out[1] = t[1];
  out[2] = t[2];
  if (s->coulomb) {
    const double M = (double)p.ng[0] * p.ng[1] * p.ng[2];
    out[3] = c1 * re;
    out[4] = c1 * qu / (2 * M);
    out[5] = s->h.slab / 2 * c1 * t[4] * t[4];
    out[6] = -c1 * p.alpha / std::sqrt(kPi) * t[5];
    out[14] = c1 * rw;
  }
  out[7] = out[3] + out[4] + out[5] + out[6];
  out[8] = t[6];
  out[9] = out[0] + out[1] + out[2] + out[7] + out[8];
  out[10] = t[3];
  out[11] = p.NN > p.N_anchor ? t[3] / (p.NN - p.N_anchor) : 0;
  out[12] = t[7];
  out[13] = t[8];
  return MD_OK;
}

extern "C" int md_energies(md_engine *e, double *out) {
  ENGINE_CHECK(e && out);
  FIRST(e);
  STATE_CHECK(e);
  return energies(e, s, P, out);
}

extern "C" int md_batch_energies(md_engine *e, int32_t slot, double *out) {
  ENGINE_CHECK(e && out);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  if (e->dirty) {
    int rc = e->rebuild();
    if (rc) return rc;
  }
  Sys *s = it->second;
  return energies(e, s, e->single[e->index_of(s)], out);
}

extern "C" int md_heights(md_engine *e, double *out) {
  ENGINE_CHECK(e && out);
  FIRST(e);
  STATE_CHECK(e);
  return heights(e, s, P, out);
}

extern "C" int md_batch_heights(md_engine *e, int32_t slot, double *out) {
  ENGINE_CHECK(e && out);
  auto it = e->slots.find(slot);
  if (it == e->slots.end()) return e->fail(MD_EINVAL, "no such slot");
  if (e->dirty) {
    int rc = e->rebuild();
    if (rc) return rc;
  }
  Sys *s = it->second;
  return heights(e, s, e->single[e->index_of(s)], out);
}

// ------------------------------------------------------- exact resume

// The complete state of a single-system engine between steps: positions,
// velocities, forces, the held Coulomb force, the LJ list with its cells
// and the control block (accumulated displacements, rebuild generation).
// Loading it into an engine made from the same parameters continues the
// trajectory bit for bit (the reference restarts from its six-decimal text
// checkpoint and rescales the velocities, main.f90:40-64).
namespace {
struct StateHead {
  char magic[8];  // "MDSTATE2"
  int64_t NN, fp32, ncell1, ccap1, nbcap1;
  int64_t Nq, ncell2, nbcap2;  // the walked reference-form real-space list (0: none)
};

// header, pos/vel/acc, held Coulomb force, LJ list lengths and slots, the
// two LJ count buffers, LJ cell members, each particle's LJ cell, [real-space
// list lengths and slots, its two count buffers,] control block
size_t state_bytes(size_t NN, size_t rs, size_t ncell1, size_t ccap1, size_t nbcap1, size_t Nq,
                   size_t ncell2, size_t nbcap2) {
  return sizeof(StateHead) + NN * 10 * 8 + 3 * NN * rs + NN * 4 + nbcap1 * NN * 4 + 2 * ncell1 * 4 +
         ncell1 * ccap1 * 4 + NN * 4 + (nbcap2 ? Nq * 4 + nbcap2 * Nq * 4 + 2 * ncell2 * 4 : 0) +
         sizeof(Ctrl);
}
size_t state_bytes(const Sys *s) {
  const bool rl = s->h.real_list != 0;
  return state_bytes(s->p.NN, s->fp32 ? 4 : 8, s->ncell1, s->h.ccap1, s->h.nbcap1, rl ? s->h.Nq : 0,
                     rl ? s->ncell2 : 0, rl ? s->h.nbcap2 : 0);
}
}  // namespace

extern "C" int md_state_size(md_engine *e, int64_t *bytes) {
  ENGINE_CHECK(e && bytes);
  FIRST(e);
  STATE_CHECK(e);
  *bytes = (int64_t)state_bytes(s);
  return MD_OK;
}

extern "C" int md_state_save(md_engine *e, void *buf, int64_t cap) {
  ENGINE_CHECK(e && buf);
  FIRST(e);
  STATE_CHECK(e);
  const size_t need = state_bytes(s);
  if ((size_t)cap < need) return e->fail(MD_EINVAL, "state buffer too small");
  const size_t NN = s->p.NN, rs = s->fp32 ? 4 : 8;
  char *o = (char *)buf;
  StateHead hd{};
  std::memcpy(hd.magic, "MDSTATE2", 8);
  hd.NN = (int64_t)NN;
  hd.fp32 = s->fp32;
  hd.ncell1 = s->ncell1;
  hd.ccap1 = s->h.ccap1;
  hd.nbcap1 = s->h.nbcap1;
  const bool rl = s->h.real_list != 0;
  hd.Nq = rl ? s->h.Nq : 0;
  hd.ncell2 = rl ? s->ncell2 : 0;
  hd.nbcap2 = rl ? s->h.nbcap2 : 0;
  std::memcpy(o, &hd, sizeof hd);
  o += sizeof hd;
  Ctrl c{};
  int rc = download_sys(e, s, (double *)o, (double *)o + 4 * NN, (double *)o + 7 * NN, &c);
  if (rc) return rc;
  o += NN * 10 * 8;
  auto get = [&](const void *src, size_t n) {
    hipError_t err = hipMemcpy(o, src, n, hipMemcpyDeviceToHost);
    o += n;
    return err;
  };
  hipError_t err = get(s->h.ccx, 3 * NN * rs);
  if (err == hipSuccess) err = get(s->h.nbn1, NN * 4);
  if (err == hipSuccess) err = get(s->h.nb1, (size_t)s->h.nbcap1 * NN * 4);
  if (err == hipSuccess) err = get(s->h.cnt1, 2 * (size_t)s->ncell1 * 4);
  if (err == hipSuccess) err = get(s->h.mem1, (size_t)s->ncell1 * s->h.ccap1 * 4);
  if (err == hipSuccess) err = get(s->h.cell1_of, NN * 4);
  if (rl) {
    if (err == hipSuccess) err = get(s->h.nbn2, (size_t)s->h.Nq * 4);
    if (err == hipSuccess) err = get(s->h.nb2, (size_t)s->h.nbcap2 * s->h.Nq * 4);
    if (err == hipSuccess) err = get(s->h.cnt2[0], (size_t)s->ncell2 * 4);
    if (err == hipSuccess) err = get(s->h.cnt2[1], (size_t)s->ncell2 * 4);
  }
  if (err != hipSuccess) return e->hip_fail(err, "state save");
  std::memcpy(o, &c, sizeof c);
  return MD_OK;
}

extern "C" int md_state_load(md_engine *e, const void *buf, int64_t len) {
  ENGINE_CHECK(e && buf);
  FIRST(e);
  StateHead hd;
  if (len < (int64_t)sizeof hd) return e->fail(MD_EINVAL, "state too short");
  std::memcpy(&hd, buf, sizeof hd);
  const bool rl = s->h.real_list != 0;
  if (std::memcmp(hd.magic, "MDSTATE2", 8) || hd.NN != s->p.NN || hd.fp32 != (int64_t)s->fp32 ||
      hd.ncell1 != s->ncell1 || hd.Nq != (rl ? s->h.Nq : 0) || hd.ncell2 != (rl ? s->ncell2 : 0))
    return e->fail(MD_EINVAL, "state of another system or precision");
  // capacities as size_lists can produce them (cell members at most 128, a
  // list at most every other particle, rounded to 8); the whole buffer is
  // checked before anything of the engine changes
  if (hd.ccap1 <= 0 || hd.ccap1 > 128 || hd.nbcap1 <= 0 || hd.nbcap1 > hd.NN + 8 ||
      (rl && (hd.nbcap2 <= 0 || hd.nbcap2 > hd.Nq + 64)) || (!rl && hd.nbcap2 != 0))
    return e->fail(MD_EINVAL, "state with invalid list capacities");
  const size_t NN = s->p.NN, rs = s->fp32 ? 4 : 8;
  if ((uint64_t)len < state_bytes(NN, rs, (size_t)hd.ncell1, (size_t)hd.ccap1, (size_t)hd.nbcap1,
                                  (size_t)hd.Nq, (size_t)hd.ncell2, (size_t)hd.nbcap2))
    return e->fail(MD_EINVAL, "state too short");
  const char *in = (const char *)buf + sizeof hd;
  int rc = upload_sys(e, s, (const double *)in, (const double *)in + 4 * NN, (const double *)in + 7 * NN);
  if (rc) return rc;
  rc = s->set_list_caps((int)hd.ccap1, (int)hd.nbcap1, (int)hd.nbcap2);
  if (rc) return e->fail(rc, s->err);
  in += NN * 10 * 8;
  auto put = [&](void *dst, size_t n) {
    hipError_t err = hipMemcpy(dst, in, n, hipMemcpyHostToDevice);
    in += n;
    return err;
  };
  hipError_t err = put(s->h.ccx, 3 * NN * rs);
  if (err == hipSuccess) err = put(s->h.nbn1, NN * 4);
  if (err == hipSuccess) err = put(s->h.nb1, (size_t)s->h.nbcap1 * NN * 4);
  if (err == hipSuccess) err = put(s->h.cnt1, 2 * (size_t)s->ncell1 * 4);
  if (err == hipSuccess) err = put(s->h.mem1, (size_t)s->ncell1 * s->h.ccap1 * 4);
  if (err == hipSuccess) err = put(s->h.cell1_of, NN * 4);
  if (rl) {
    if (err == hipSuccess) err = put(s->h.nbn2, (size_t)s->h.Nq * 4);
    if (err == hipSuccess) err = put(s->h.nb2, (size_t)s->h.nbcap2 * s->h.Nq * 4);
    if (err == hipSuccess) err = put(s->h.cnt2[0], (size_t)s->ncell2 * 4);
    if (err == hipSuccess) err = put(s->h.cnt2[1], (size_t)s->ncell2 * 4);
  }
  if (err == hipSuccess) err = hipMemcpy(s->h.ctrl, in, sizeof(Ctrl), hipMemcpyHostToDevice);
  s->step_dev = -1;
  if (err == hipSuccess) err = null_stream_done();
  if (err != hipSuccess) return e->hip_fail(err, "state load");
  return e->rebuild();
}
