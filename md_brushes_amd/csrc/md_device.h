// md_device.h -- device-side data layout of the MI355X brush engine.
//
// One engine holds R replicas ("systems") of possibly different sizes.  Every
// kernel covers all replicas in one launch: its blocks are split into
// per-replica ranges by a prefix table (blk[r] .. blk[r+1]) so that every
// block belongs to exactly one replica and reads that replica's descriptor
// with uniform (scalar) loads.  A single system is the case R = 1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdk {

constexpr int kBlock = 256;
constexpr int kMaxOrder = 8;
constexpr int kTile = 16;         // spreading tile edge (mesh points)
constexpr int kTileSub = 8;       // sub-lists per spreading tile (spreads the appends)
constexpr int kPlaneBlock = 512;  // threads of the fused FFT plane kernels
constexpr int kPlaneSpreadPer = 16;  // plane points per thread of k_plane_spread (K1 K2 <= 16 kPlaneBlock)
constexpr int kPlaneSpreadTable = 272;  // LDS bytes of k_plane_spread's sub-list table, after the twiddles
constexpr int kSRLanes = 4;       // lanes per particle in the LJ pass of the short-range kernel
constexpr int kSRBlock = 512;     // its workgroup (LJ on 4 wavefronts, the per-particle
constexpr int kSRParticles = 64;  // parts on the others) takes 64 particles
constexpr int kRealChunk = 4;     // target charges per wavefront work item of the real-space sweep
constexpr int kRealChunkWG = 16;  // ... per workgroup work item (replicas with dense cells)
constexpr int kGatherLanes = 4;   // lanes per charge of the reciprocal-force gather
constexpr int kRealRound = 256;   // real-space candidates tested per compaction round (4 per lane)
constexpr int kLJStage = 256;     // LJ fill pass: candidates staged in order per target
constexpr int kHistBins = 500;  // SizeHist of the z-density profile (glob_var.f90:59)
constexpr int kMaxFac = 24;     // FFT passes of a mesh length that is not a power of two
constexpr int kXCD = 8;         // MI355X accelerator dies: workgroup b of a launch runs on die b % 8

// Per-replica control block, written by kernels.
struct Ctrl {
  double dr_max1, dr_max2;        // accumulated (compt_acc.f90:16,41)
  double pend_d1, pend_d2;        // a replayed step's maximum displacements, applied
                                  // to dr_max1/2 by the step's list kernel (pend != 0)
  int pend;
  int flag1, flag2;               // rebuild LJ / real list after this step
  int gen1, gen2;                 // rebuild generations (count-buffer parity)
  int err;                        // first error code (MD_E*)
  int err_info;                   // bond or particle index of the error
  int err_step;
  int ovf;                        // which capacity overflowed (bitmask)
  int max_nb1, max_nb2;           // largest list length seen at last build
  int max_cell1, max_cell2;       // largest cell occupancy seen at last build
  int step;                       // step counter of graph-replayed steps
};

// Per-replica descriptor (read-only in kernels).
struct Rep {
  // sizes
  int NN, Nq, N_anchor, real_verlet;
  // the real-space force walks the reference-form list (real_verlet and
  // fewer than three list cells along x or y, where the reference's stencil
  // lists a neighbour cell's pairs more than once); otherwise the cell sweep
  int real_list;
  int Npe;                         // polymer beads are particles 0..Npe-1
  int Nga, arm, Nma, Ngl, Nml;     // stars (first), then linear chains
  int preheat_end;                 // batch replicas: steps 1..preheat_end rescale
                                   // velocities every 20 steps (0: none)
  // geometry
  double Lx, Ly, Lz, hLx, hLy;
  // integrator (init_update.f90:555-607)
  double dt, c1m, c1p, sig;       // c1m = 1-xi*dt/2, c1p = 1+xi*dt/2
  double Beta;
  double EF;
  uint64_t seed;
  // short range (compt_acc.f90:895-977, 865-892)
  double rcl2, rcl3, rcl4, sigma2, sigma3, sigma_s;
  double wall_r, wall_rs, wall_c;  // rcl/5^(1/6), rcl_s/5^(1/6), 0.4 (default real)
  double kR02, R0_2;               // kFENE*R0_2, R0_2
  // Coulomb (compt_acc.f90:980-1049, 1213-1246)
  double c1, alpha, alpha2, c2, rcc2, slab;  // lb/Beta, 2a/sqrt(pi), 4pi/V
  int K[3], order;                 // mesh; the spline order when equal on all axes, else 0
  int ord[3];                      // spline order per axis, ordr(1:3) (compt_acc.f90:1232-1234)
  double kg[3];                    // K_d / gdim_d
  double fnorm[3];                 // K_d/gdim_d/(K1 K2 K3)
  int Kh;                          // K1/2+1
  int lK[3];                       // log2 K_d (power-of-two axes)
  int pw2[3];                      // 1: K_d is a power of two (radix 8/4/2 passes)
  int nfac[3];                     // other lengths: one Stockham pass per factor
  int fac[3][kMaxFac];             // (radix 2, 3, 4, 5, 8 butterflies; any other prime by a DFT)
  int px_tile, py_tile, pz_tile;   // FFT tile sizes (row pairs, lines, lines)
  // Verlet lists (compt_acc.f90:185-314, 491-626)
  double rv1sq, rv2sq;
  int nc1[3], nc2[3];
  double rcel1[3], rcel2[3];
  int ccap1, ccap2;                // cell capacity
  int nbcap1, nbcap2;              // neighbour capacity
  double skin1, skin2;             // (rvl-rcl_s)/2, rsk/2
  // --- device pointers ---
  double4 *pos;                    // (x,y,z,q) FP64, canonical
  float4 *pos32;                   // FP32 copy for the FP32 force path
  // the other position buffer: replayed steps read pos and write the next
  // step's drifted positions here (two descriptor tables swap the pair)
  double4 *pos_nx;
  float4 *pos32_nx;
  int par;                         // which buffer pos is (0 / 1)
  // max |dr| of a step (all, charged) per group of 64 consecutive particles,
  // one slot of ngrp groups per position buffer: the drift that produces
  // buffer q's positions writes slot q (plain stores, every group)
  double2 *dpart;
  int ngrp;
  double *vx, *vy, *vz;
  double *ax, *ay, *az;
  void *ccx, *ccy, *ccz;           // held Coulomb force, real type of the path
  int *flags;                      // bit0 anchor, bit1 charged
  int *charge;                     // Nq: particle index of charge m
  int *bstart;                     // NN+1 CSR of bonded partners
  int *bpart;
  int4 *bond4;                     // each particle's first four partners, bond order (-1: none)
  // LJ cells: two count buffers used in turn (gen1 & 1 selects the one a
  // build fills; it clears the other at its particles' previous cells),
  // members as binned, each particle's cell (-1: none)
  int *cnt1, *mem1, *cell1_of;
  int *cnt2[2], *mem2;             // cells of the reference-form real-space list
  // real-space cell sweep at each Coulomb evaluation (edge >= r_c)
  int real_cells;                  // 1: the real-space cell sweep runs (charged systems)
  int real_wg;                     // sweep items taken by workgroups (dense cells)
  int nc3[3];
  double rcel3[3];
  int ccap3;
  int *cnt3, *mem3, *scnt3;        // mem3: the cells' charges as particle indices
  void *pos3s;                     // positions in cell-sorted order (float4 / double4)
  void *freal;                     // real-space pair sums per particle, 3 x NN (real type;
                                   // charged particles only)
  int *nbn1, *nb1;                 // LJ list: length per particle, ELL slots
  int *nbn2, *nb2;                 // real list: per charge slot m
  void *meshr;                     // charge mesh Q (real type of the path) K3*K2*K1
  // spreading tiles: each tile lists the charges with spline points in it
  int tsz[3], trem[3], ntl[3];     // tiles per axis: the first trem of edge tsz + 1, then tsz
  int capt;                        // capacity of each of a tile's kTileSub sub-lists
  int *tcnt, *tmem;                // counts; particle indices (list, sub, slot): a list per
                                   // spreading tile, or per first z index (plane spreading)
  int nlist;                       // counts to clear after a gather: max(tiles, K3) x kTileSub
  int plane_fused;                 // x and y FFT passes fused per z-plane
  void *spec;                      // half spectrum K3*K2*Kh complex
  void *meshu;                     // real potential mesh K3*K2*K1
  void *bc;                        // influence function K3*K2*Kh real
  const void *tw[3];               // twiddles per axis (K_d/2 complex)
  double *fsp;                     // 3 x NN reciprocal force before mean removal (charged
                                   // particles only)
  double *part;                    // block partial sums: 4 per charge block
  int nqblk;                       // number of charge blocks
  Ctrl *ctrl;
};

// Spreading tiles along axis d: ntl tiles, the first trem of edge tsz + 1
// and the rest of edge tsz (every edge >= the spline order, at most kTile).
__host__ __device__ inline int tile_start(const Rep &S, int d, int t) {
  return t * S.tsz[d] + (t < S.trem[d] ? t : S.trem[d]);
}
__host__ __device__ inline int tile_edge(const Rep &S, int d, int t) {
  return S.tsz[d] + (t < S.trem[d] ? 1 : 0);
}
__host__ __device__ inline int tile_of(const Rep &S, int d, int b) {
  const int big = S.trem[d] * (S.tsz[d] + 1);
  return b < big ? b / (S.tsz[d] + 1) : S.trem[d] + (b - big) / S.tsz[d];
}

// Work map: block ranges per replica.
struct WMap {
  const int *blk;  // nrep+1 prefix
  int nrep;
  int total;
};

__device__ __forceinline__ int find_rep(const WMap &m, int b) {
  int lo = 0, hi = m.nrep;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (m.blk[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// Fixed-point scale of the spread mesh: 2^40.
constexpr double kQScale = 1099511627776.0;
constexpr double kQInv = 1.0 / 1099511627776.0;

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                              uint32_t c3, uint32_t k0, uint32_t k1,
                                              uint32_t out[4]) {
#pragma unroll
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

constexpr uint32_t kNoiseTag = 0x4D444E53u;  // counter word 2 of the noise stream
constexpr uint32_t kVelTag = 0x4D445645u;    // initial velocities
constexpr uint32_t kInitTag = 0x4D44494Eu;   // initial configuration

}  // namespace mdk
