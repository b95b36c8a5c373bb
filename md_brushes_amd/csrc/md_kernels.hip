// md_kernels.hip -- gfx950 kernels of the brush engine.
//
// Every kernel covers all replicas of an engine in one launch (block ranges
// per replica, see WMap).  R is the real type of the force path: double for
// the plain FP64 path, float for the FP32 path (positions, velocities, the
// integrator, M_z and all reductions stay double).
#include "md_kernels.h"

#include <type_traits>

#include "../../include/md_engine.h"

namespace mdk {

template <typename R> struct CT;
template <> struct CT<float> { using C = float2; };
template <> struct CT<double> { using C = double2; };

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }

__device__ __forceinline__ void set_error(Ctrl *c, int code, int info, int step) {
  if (atomicCAS(&c->err, 0, code) == 0) {
    c->err_info = info;
    c->err_step = step;
  }
}

// Slot of this lane's entry in the list counted by *ctr (null: no entry).
// Runs of adjacent lanes with the same counter (neighbouring particles along
// a chain mostly share their cell) take one atomic per run; the runs' atomics
// are issued together.  Every lane of the wavefront calls it.
// A lane's run of equal counters in its wavefront: the run's head lane h,
// and, at a head, the run's length n.
struct WaveRun {
  int h, n;
  bool head;
};
__device__ __forceinline__ WaveRun wave_run(const int *ctr) {
  const int lane = threadIdx.x & 63;
  const unsigned long long me = (unsigned long long)ctr;
  const unsigned long long prev = __shfl_up(me, 1, 64);
  const bool head = lane == 0 || prev != me;
  const unsigned long long heads = __ballot(head);
  const unsigned long long le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  WaveRun w;
  w.h = 63 - __clzll((long long)(heads & le));  // head of this lane's run
  const unsigned long long after = heads & ~le;
  w.n = (after ? __ffsll((long long)after) - 1 : 64) - lane;  // to the first head past this lane
  w.head = head;
  return w;
}
__device__ __forceinline__ int wave_append(int *ctr) {
  const WaveRun w = wave_run(ctr);
  int base = 0;
  if (w.head && ctr) base = atomicAdd(ctr, w.n);
  base = __shfl(base, w.h, 64);
  return ctr ? base + ((int)(threadIdx.x & 63) - w.h) : -1;
}

// LJ rebuild queue: the replicas whose list is rebuilt after this step and
// the range [start, start + NN) each takes in the global particle index the
// list kernels stride over.  One 64-bit head packs (entries << 40) | total,
// so that one atomic hands out the entry and its range in the same order;
// a spare 64-bit word, then rep[cap] and start[cap].  A rep word also
// carries the cell-count buffer (bit 30) the replica's particles are binned
// into, so that the list kernels need not read the rebuild generation.
constexpr int kQShift = 40;
constexpr int kQBuf = 1 << 30;
__device__ __forceinline__ int *lq_rep(unsigned long long *lq) { return (int *)(lq + 2); }
__device__ __forceinline__ int *lq_start(unsigned long long *lq, int cap) {
  return (int *)(lq + 2) + cap;
}
__device__ __forceinline__ void lq_push(unsigned long long *lq, int cap, int r, int nn, int buf) {
  const unsigned long long old = atomicAdd(lq, (1ull << kQShift) + (unsigned long long)nn);
  const int e = (int)(old >> kQShift);
  if (e < cap) {
    lq_rep(lq)[e] = r | (buf ? kQBuf : 0);
    lq_start(lq, cap)[e] = (int)(old & ((1ull << kQShift) - 1));
  }
}
__device__ __forceinline__ int lq_entry_rep(int w) { return w & (kQBuf - 1); }
__device__ __forceinline__ int lq_entry_buf(int w) { return (w & kQBuf) ? 1 : 0; }
// Entry of global particle index g: the largest e with start[e] <= g.
__device__ __forceinline__ int lq_find(const int *start, int n, int g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// x or y minimum image of rij_and_rr (glob_var.f90:72-105): d - L above L/2,
// d + L at or below -L/2, as selects (no divergent branch in a wavefront).
template <typename R> __device__ __forceinline__ R image(R d, R L, R hL) {
  const R lo = d - L, hi = d + L;
  return d > hL ? lo : (d <= -hL ? hi : d);
}

// Squared distance with the x,y minimum image of rij_and_rr
// (glob_var.f90:72-105), rounded term by term as the reference writes it.
__device__ __forceinline__ double rsq_image(double4 a, double4 b, double Lx, double Ly,
                                            double hLx, double hLy, double &dx, double &dy,
                                            double &dz) {
#pragma clang fp contract(off)
  dx = a.x - b.x;
  dy = a.y - b.y;
  dz = a.z - b.z;
  const double xl = dx - Lx, xh = dx + Lx, yl = dy - Ly, yh = dy + Ly;
  dx = dx > hLx ? xl : (dx <= -hLx ? xh : dx);
  dy = dy > hLy ? yl : (dy <= -hLy ? yh : dy);
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Value of v in lane l (l uniform over the wavefront).
__device__ __forceinline__ float lane_value(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Lane moves of the DPP modifier (quad permutes, row mirrors, row
// broadcasts); lanes of rows outside row_mask RM read 0.
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_move(float v) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, RM, 0xf, false));
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, RM, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, RM, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sum over the 64 lanes in a fixed order, every lane taking the total: pair
// and quad sums (quad permutes), the 8 and 16 lanes of a row (half-row and
// row mirrors), then rows 0+1 and 2+3 (lane 15 broadcast into rows 1, 3)
// and all four (lane 31 broadcast into rows 2, 3); lane 63 holds the total.
// DPP moves stay in the vector ALU, where a shuffle takes an LDS round trip.
template <typename T> __device__ __forceinline__ T wave_total(T v) {
  v += dpp_move<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xf>(v);  // row_mirror
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31
  return lane_value(v, 63);
}

// Maximum over the 64 lanes of non-negative values (displacements), every
// lane taking it: the DPP moves of wave_total with fmax (lanes a move
// leaves out read 0, which no maximum of non-negative values takes; fmax
// passes over NaN).  The maximum does not depend on the order of the
// moves.
__device__ __forceinline__ double wave_max_nonneg(double v) {
  v = fmax(v, dpp_move<0xb1, 0xf>(v));
  v = fmax(v, dpp_move<0x4e, 0xf>(v));
  v = fmax(v, dpp_move<0x141, 0xf>(v));
  v = fmax(v, dpp_move<0x140, 0xf>(v));
  v = fmax(v, dpp_move<0x142, 0xa>(v));
  v = fmax(v, dpp_move<0x143, 0xc>(v));
  return lane_value(v, 63);
}

// Orders LDS accesses of the lanes of one wavefront (LDS operations of a
// wavefront complete in order; the fences keep the compiler from moving
// memory accesses across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads one word of each 64-byte line of a replica's descriptor with scalar
// loads issued together, waiting once: the scalar data cache then holds the
// whole descriptor, and the kernel's later reads of its fields (which the
// compiler issues where they are used, each a wait of its own) hit it
// instead of going to memory one after another.
__device__ __forceinline__ void desc_prefetch(const Rep &S) {
  static_assert(sizeof(Rep) <= 24 * 64, "the descriptor spans at most 24 lines");
  constexpr int L = (int)((sizeof(Rep) + 63) / 64);
  const unsigned *q = (const unsigned *)&S;
  unsigned v[24];
#pragma unroll
  for (int k = 0; k < 24; k++) v[k] = k < L ? q[16 * k] : 0u;
  // every word an operand of one statement: the loads are all in flight
  // before the one wait
  asm volatile("" ::"s"(v[0]), "s"(v[1]), "s"(v[2]), "s"(v[3]), "s"(v[4]), "s"(v[5]), "s"(v[6]),
               "s"(v[7]), "s"(v[8]), "s"(v[9]), "s"(v[10]), "s"(v[11]), "s"(v[12]), "s"(v[13]),
               "s"(v[14]), "s"(v[15]), "s"(v[16]), "s"(v[17]), "s"(v[18]), "s"(v[19]), "s"(v[20]),
               "s"(v[21]), "s"(v[22]), "s"(v[23]));
}

// ------------------------------------------------------------- integrator

// First half kick of new_position (init_update.f90:581-589), mass 1.
template <typename K>
__device__ __forceinline__ double half_kick(const K &S, double a, double v) {
#pragma clang fp contract(off)
  return a * S.dt / 2 + S.c1m * v;
}

// Drift and period_condition_pos (init_update.f90:594-601, 631-642): x += v dt,
// x and y wrapped to (-L/2, L/2]; returns |v dt|.
template <typename K>
__device__ __forceinline__ double drift(const K &S, double4 &p, double vx, double vy, double vz) {
#pragma clang fp contract(off)
  const double dt = S.dt;
  const double dx = vx * dt, dy = vy * dt, dz = vz * dt;
  p.x = p.x + dx;
  p.y = p.y + dy;
  p.z = p.z + dz;
  const double dr = sqrt(dx * dx + dy * dy + dz * dz);
  if (p.x > S.hLx) p.x = p.x - S.Lx;
  else if (p.x <= -S.hLx) p.x = p.x + S.Lx;
  if (p.y > S.hLy) p.y = p.y - S.Ly;
  else if (p.y <= -S.hLy) p.y = p.y + S.Ly;
  return dr;
}

// Maximum displacement of one wavefront's 64 consecutive particles (group g)
// into slot q of the group maxima; every lane calls it.
template <typename K>
__device__ __forceinline__ void dmax_group(const K &S, int q, int g, double d1, double d2) {
  d1 = wave_max_nonneg(d1);
  d2 = wave_max_nonneg(d2);
  if ((threadIdx.x & 63) == 0 && g < S.ngrp) S.dpart[(size_t)q * S.ngrp + g] = make_double2(d1, d2);
}

// First half of new_position (init_update.f90:581-601) with
// period_condition_pos: half kick, drift, wrap, maximum displacement of this
// step per particle group.  store_v = 0 keeps v: the step's short-range
// kernel then recomputes the kicked velocity (the form that starts a run of
// replayed steps, which drift inside the short-range kernel).  Also empties
// this step's LJ rebuild queue.
__global__ __launch_bounds__(kBlock) void k_kick_drift(const Rep *__restrict__ reps, WMap wm,
                                                       unsigned long long *lq, int store_v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *lq = 0ull;
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int blk = blockIdx.x - wm.blk[r];
  const int i = blk * kBlock + threadIdx.x;
  double d1 = 0, d2 = 0;
  if (i < S.NN) {
    const double vx = half_kick(S, S.ax[i], S.vx[i]);
    const double vy = half_kick(S, S.ay[i], S.vy[i]);
    const double vz = half_kick(S, S.az[i], S.vz[i]);
    double4 p = S.pos[i];
    const double dr = drift(S, p, vx, vy, vz);
    if (store_v) {
      S.vx[i] = vx;
      S.vy[i] = vy;
      S.vz[i] = vz;
    }
    S.pos[i] = p;
    if (S.pos32) S.pos32[i] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
    d1 = dr;
    if (p.w != 0) d2 = dr;
  }
  dmax_group(S, S.par, blk * (kBlock / 64) + (threadIdx.x >> 6), d1, d2);
}

// --------------------------------------------------------- short range

// Wall force on a particle (compt_acc.f90:956-975).
__device__ __forceinline__ double wall_force(const Rep &S, double4 pi) {
  double wz = 0;
  const double z = pi.z, Lz = S.Lz;
  if (!(fabs(z) < 0.01 || fabs(z - Lz) < 0.01)) {
    if (pi.w != 3) {
      if (z < S.wall_r) {
        double dl = z, d5 = dl * dl * dl * dl * dl, d11 = d5 * d5 * dl;
        wz = S.wall_c * (3 / d11 - 1 / d5) * dl;
      } else if (Lz - z < S.wall_r) {
        double dl = Lz - z, d5 = dl * dl * dl * dl * dl, d11 = d5 * d5 * dl;
        wz = -S.wall_c * (3 / d11 - 1 / d5) * dl;
      }
    } else {
      const double ss = S.sigma_s;
      if (z < S.wall_rs) {
        double dl = z, t = ss / dl, t5 = t * t * t * t * t, t11 = t5 * t5 * t;
        wz = S.wall_c * (3 * t11 - t5) * dl / ss / ss;
      } else if (Lz - z < S.wall_rs) {
        double dl = Lz - z, t = ss / dl, t5 = t * t * t * t * t, t11 = t5 * t5 * t;
        wz = -S.wall_c * (3 * t11 - t5) * dl / ss / ss;
      }
    }
  }
  return wz;
}

// Three standard normals of particle i at `step`: one Philox4x32-10 block,
// counter (i, step, tag, 0), key = seed; words mapped to (0,1) as
// (w + 0.5) / 2^32; Box-Muller, both branches of the first pair and the
// cosine branch of the second (the reference's gauss_dist,
// glob_var.f90:108-121, draws from an OS-seeded stream instead).
__device__ __forceinline__ void noise_normals(const Rep &S, int i, int step, double &n0, double &n1,
                                              double &n2) {
  uint32_t w[4];
  philox4x32_10((uint32_t)i, (uint32_t)step, kNoiseTag, 0u, (uint32_t)S.seed,
                (uint32_t)(S.seed >> 32), w);
  const double s32 = 1.0 / 4294967296.0;
  const double u0 = ((double)w[0] + 0.5) * s32, u1 = ((double)w[1] + 0.5) * s32;
  const double u2 = ((double)w[2] + 0.5) * s32, u3 = ((double)w[3] + 0.5) * s32;
  const double r1 = sqrt(-2 * log(u0)), r2 = sqrt(-2 * log(u2));
  double sn, cs;
  sincospi(2 * u1, &sn, &cs);
  n0 = r1 * cs;
  n1 = r1 * sn;
  n2 = r2 * cospi(2 * u3);
}

__device__ __forceinline__ void bin_lj_particle(const Rep *Sp, int i, int buf);

// What the state wavefront of k_short_range uses after the workgroup barrier:
// the replica's step constants and pointers, read from the descriptor
// before the barrier (a descriptor field first read after it costs that
// wavefront a memory round trip of its own, on the kernel's critical path;
// hold() keeps the compiler from moving the loads past the barrier).
struct StepK {
  double dt, c1m, c1p, sig, EF, Lx, Ly, hLx, hLy;
  double *vx, *vy, *vz, *ax, *ay, *az;
  double4 *pos_nx;
  float4 *pos32_nx;
  double2 *dpart;
  Ctrl *ctrl;
  int NN, ngrp, par;
  __device__ __forceinline__ void hold() const {
    asm volatile("" ::"s"(dt), "s"(c1m), "s"(c1p), "s"(sig), "s"(EF), "s"(Lx), "s"(Ly), "s"(hLx),
                 "s"(hLy));
    asm volatile("" ::"s"(vx), "s"(vy), "s"(vz), "s"(ax), "s"(ay), "s"(az), "s"(pos_nx),
                 "s"(pos32_nx), "s"(dpart), "s"(ctrl));
    asm volatile("" ::"s"(NN), "s"(ngrp), "s"(par));
  }
};

__device__ __forceinline__ StepK step_constants(const Rep &S) {
  StepK k;
  k.dt = S.dt;
  k.c1m = S.c1m;
  k.c1p = S.c1p;
  k.sig = S.sig;
  k.EF = S.EF;
  k.Lx = S.Lx;
  k.Ly = S.Ly;
  k.hLx = S.hLx;
  k.hLy = S.hLy;
  k.vx = S.vx;
  k.vy = S.vy;
  k.vz = S.vz;
  k.ax = S.ax;
  k.ay = S.ay;
  k.az = S.az;
  k.pos_nx = S.pos_nx;
  k.pos32_nx = S.pos32_nx;
  k.dpart = S.dpart;
  k.ctrl = S.ctrl;
  k.NN = S.NN;
  k.ngrp = S.ngrp;
  k.par = S.par;
  return k;
}

// The sums behind a Coulomb evaluation's store, by one wavefront: the
// gather's block partials of the reciprocal force (3) and of q z lb/Beta
// (M_z), lane l summing blocks l, l + 64, ... in order, then a fixed DPP
// tree over the lanes (wave_total).  k_coul_store and the short-range
// kernel of a Coulomb step (SR_COUL) both take their totals here, so the
// held force is the same bits whichever stores it.
template <int U>
__device__ __forceinline__ void coul_totals(const Rep &S, double tot[4]) {
  const int lane = threadIdx.x & 63;
  const double4 *pb = (const double4 *)S.part;
  double v[4] = {0, 0, 0, 0};
  // U blocks per lane at a time, their loads issued together (one memory
  // round trip per 64 U blocks), added in block order
  for (int b0 = lane; b0 < S.nqblk; b0 += 64 * U) {
    double4 x[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b = b0 + 64 * u;
      x[u] = b < S.nqblk ? pb[b] : make_double4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (b0 + 64 * u < S.nqblk) {
        v[0] += x[u].x;
        v[1] += x[u].y;
        v[2] += x[u].z;
        v[3] += x[u].w;
      }
  }
#pragma unroll
  for (int d = 0; d < 4; d++) tot[d] = wave_total(v[d]);
}

// acc_c of one charge: the reciprocal force with its mean removed, scaled by
// lb/Beta (compt_acc.f90:1242-1244), plus the real-space pair force (scaled
// here), minus the slab correction (compt_acc.f90:1045-1048).  Rounded
// term by term (no contraction), so that every kernel that inlines it gets
// the same bits.
// A charged particle's entries of the two halves (reciprocal, real space).
template <typename R> struct CoulParts {
  double fs[3];
  R fr[3];
  __device__ __forceinline__ void load(const Rep &S, int i) {
    const R *f = (const R *)S.freal;
    const size_t NN = S.NN;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      fs[d] = S.fsp[d * NN + i];
      fr[d] = f[d * NN + i];
    }
  }
};

template <typename R>
__device__ __forceinline__ void coul_value(const Rep &S, const CoulParts<R> &c, double q,
                                           const double *tot, double &ax, double &ay, double &az) {
#pragma clang fp contract(off)
  const double c1 = S.c1, inv = 1.0 / S.Nq;
  ax = (c.fs[0] - tot[0] * inv) * c1 + (double)c.fr[0] * c1;
  ay = (c.fs[1] - tot[1] * inv) * c1 + (double)c.fr[1] * c1;
  az = (c.fs[2] - tot[2] * inv) * c1 + (double)c.fr[2] * c1 - S.slab * q * tot[3];
}

// Compute_Force without the Coulomb evaluation (compt_acc.f90:847-861):
// held Coulomb force + lj_force (compt_acc.f90:895-977) + fene_force
// (compt_acc.f90:865-892) + noise + field, anchors zeroed (under SR_COUL
// the held force is first assembled from the evaluation just done, as
// k_coul_store would); optionally the second half kick of new_position
// (init_update.f90:606), for replayed steps (SR_DRIFT) the next step's half
// kick and drift into the other position buffer (init_update.f90:581-601,
// 631-642), and the dr_max
// bookkeeping of new_position / update_verlet_list (init_update.f90:600-601,
// compt_acc.f90:717-724).  A workgroup takes 64 particles and gives its
// wavefronts the independent parts, so that their memory latencies overlap:
// wavefronts 0-3 sum the LJ lists (kSRLanes lanes per particle, lane g
// takes slots g, g + kSRLanes, ..., a fixed tree over the lanes), wavefront
// 4 draws the noise, 5 the wall and FENE forces (bond order), 6 loads the
// particle's state and then adds everything up, kicks, drifts and stores;
// wavefront 7 takes the rebuild decision (the replica's first workgroup, or
// under SR_BIN every workgroup, which then bins its own particles).
template <typename R, bool COUL>
__global__ __launch_bounds__(kSRBlock) void k_short_range(const Rep *__restrict__ reps, WMap wm,
                                                          int mode, int step_arg,
                                                          const double *__restrict__ eta,
                                                          double *__restrict__ out,
                                                          unsigned long long *lq, int lq_cap,
                                                          unsigned long long *lq_nx) {
  constexpr int G = kSRLanes, NP = kSRParticles;
  static_assert(NP == 64 && G * NP == 256 && kSRBlock == 512,
                "LJ on waves 0-3, one particle per lane on the others; groups of 64 for dmax_group");
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int blk = blockIdx.x - wm.blk[r];
  const int step = step_arg >= 0 ? step_arg : S.ctrl->step;
  // the wavefront index as a uniform value: each wavefront's part is then a
  // scalar branch, and what wave 6 keeps across the barrier stays in scalar
  // registers where it is uniform (70 VGPRs instead of 88: three
  // workgroups per CU)
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int i = blk * NP + lane;  // the particle of this lane on waves 4..7
  const bool live = i < S.NN;
  __shared__ R lj[3][NP], fb[3][NP];
  __shared__ double wzs[NP], nzs[3][NP];
  __shared__ double ctot[4];  // SR_COUL: the Coulomb evaluation's totals (coul_totals)
  // wave 6's registers across the barrier: assigned on its path only (each
  // under the mode bit that later reads it), so that no other wavefront's
  // branch holds them live
  double4 pi;
  double vx, vy, vz, ax0, ay0, az0, cx, cy, cz;
  int anchor;
  bool charged;
  CoulParts<R> cp;
  StepK K;
  // SR_COUL (the Coulomb step's launch, its own instantiation, so that the
  // other steps' kernel carries none of its code or registers); uniform
  const bool coul = COUL && (mode & SR_COUL) && S.nqblk > 0;
  if (w < 4) {  // WCA-LJ over the full list: class by |q_i|+|q_j| (compt_acc.f90:922-952)
    const int g = threadIdx.x & (G - 1), pl = threadIdx.x / G;
    const int ip = blk * NP + pl;
    const bool lv = ip < S.NN;
    const int n = lv ? min(S.nbn1[ip], S.nbcap1) : 0;
    const R hLx = (R)S.hLx, hLy = (R)S.hLy, Lx = (R)S.Lx, Ly = (R)S.Ly;
    const R rcl2 = (R)S.rcl2, rcl3 = (R)S.rcl3, rcl4 = (R)S.rcl4;
    const R sig2 = (R)S.sigma2, sig3 = (R)S.sigma3;
    R xi = 0, yi = 0, zi = 0, qa = 0;
    if (lv) {
      if constexpr (sizeof(R) == 4) {
        float4 p = S.pos32[ip];
        xi = p.x; yi = p.y; zi = p.z; qa = fabsf(p.w);
      } else {
        double4 p = S.pos[ip];
        xi = p.x; yi = p.y; zi = p.z; qa = fabs(p.w);
      }
    }
    R fx = 0, fy = 0, fz = 0;
    const int *nb = S.nb1 + ip;
    using V = typename std::conditional<sizeof(R) == 4, float4, double4>::type;
    const V *pv = (const V *)(sizeof(R) == 4 ? (const void *)S.pos32 : (const void *)S.pos);
    auto pair = [&](const V pj) {
      R dx = xi - (R)pj.x, dy = yi - (R)pj.y, dz = zi - (R)pj.z;
      dx = image(dx, Lx, hLx);
      dy = image(dy, Ly, hLy);
      const R rsq = dx * dx + dy * dy + dz * dz;
      const int m = (int)(qa + fabs((R)pj.w));  // |q| are small integers: exact
      R cut, s2;
      if (m == 6) { cut = rcl3; s2 = sig2; }
      else if (m < 3) { cut = rcl2; s2 = (R)1; }
      else { cut = rcl4; s2 = sig3; }
      if (rsq < cut) {
        R gg;
        if constexpr (sizeof(R) == 4) {  // one v_rcp_f32 (about 1 ulp) for both divisions by r^2
          const float ir = __builtin_amdgcn_rcpf(rsq), ir2 = s2 * ir, ir6 = ir2 * ir2 * ir2;
          gg = ir6 * (ir6 - 0.5f) * ir;
        } else {
          const R ir2 = s2 / rsq, ir6 = ir2 * ir2 * ir2;
          gg = ir6 * (ir6 - (R)0.5) / rsq;
        }
        fx += gg * dx;
        fy += gg * dy;
        fz += gg * dz;
      }
    };
    // slots g, g + G, ... in order; four at a time, their index and position
    // loads issued together (the same sums as one at a time).  The first
    // four slots' indices are loaded beside the list length (slots past the
    // length are within the list's capacity and left unused), so a short
    // list costs two dependent loads: indices, then positions.
    {
      int jp[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int k = g + u * G;
        jp[u] = lv && k < S.nbcap1 ? nb[(size_t)k * S.NN] : 0;
      }
      int j[4];
#pragma unroll
      for (int u = 0; u < 4; u++) j[u] = g + u * G < n ? jp[u] : -1;
      V pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (j[u] >= 0) pj[u] = pv[j[u]];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (j[u] >= 0) pair(pj[u]);
    }
    for (int k = g + 4 * G; k < n; k += 4 * G) {
      int j[4];
#pragma unroll
      for (int u = 0; u < 4; u++) j[u] = k + u * G < n ? nb[(size_t)(k + u * G) * S.NN] : -1;
      V pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (j[u] >= 0) pj[u] = pv[j[u]];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (j[u] >= 0) pair(pj[u]);
    }
    static_assert(G == 4, "the lanes of a particle are a quad of the DPP tree");
    fx += dpp_move<0xb1, 0xf>(fx);  // fixed tree over the particle's lanes: xor 1, then xor 2
    fy += dpp_move<0xb1, 0xf>(fy);
    fz += dpp_move<0xb1, 0xf>(fz);
    fx += dpp_move<0x4e, 0xf>(fx);
    fy += dpp_move<0x4e, 0xf>(fy);
    fz += dpp_move<0x4e, 0xf>(fz);
    if (g == 0) {
      lj[0][pl] = fx * (R)48;
      lj[1][pl] = fy * (R)48;
      lj[2][pl] = fz * (R)48;
    }
  } else if (w == 4) {  // noise
    if (live && (mode & (SR_NOISE_ETA | SR_NOISE_PHILOX))) {
      double n0, n1, n2;
      if (mode & SR_NOISE_ETA) {
        n0 = eta[i];
        n1 = eta[(size_t)S.NN + i];
        n2 = eta[2 * (size_t)S.NN + i];
      } else {
        noise_normals(S, i, step, n0, n1, n2);
      }
      nzs[0][lane] = n0;
      nzs[1][lane] = n1;
      nzs[2][lane] = n2;
    }
    if (coul) {  // the totals for wavefront 6, after the barrier
      double t[4];
      coul_totals<4>(S, t);
      if (lane == 0)
        for (int d = 0; d < 4; d++) ctot[d] = t[d];
    }
  } else if (w == 5) {  // wall and FENE over the bonded partners of i
    if (live) {
      // the first four partners (bond order) come with the position; a
      // fifth and more (branch points of stars with five or more arms) from
      // the CSR lists
      const double4 p = S.pos[i];
      const int4 b4 = S.bond4[i];
      const R R02 = (R)S.R0_2, kR = (R)S.kR02;
      R fx = 0, fy = 0, fz = 0;
      const int b1 = b4.w >= 0 ? S.bstart[i + 1] : 0;  // past the fourth: CSR (else no load)
      // bonds in order, four at a time: the partner and position loads issued together
      for (int b = 0; b == 0 || (b1 > 0 && S.bstart[i] + b < b1); b += 4) {
        int nn[4];
        if (b == 0) {
          nn[0] = b4.x;
          nn[1] = b4.y;
          nn[2] = b4.z;
          nn[3] = b4.w;
        } else {
          const int b0 = S.bstart[i] + b;
#pragma unroll
          for (int u = 0; u < 4; u++) nn[u] = b0 + u < b1 ? S.bpart[b0 + u] : -1;
        }
        double4 pj[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (nn[u] >= 0) pj[u] = S.pos[nn[u]];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (nn[u] < 0) continue;
          double dxd, dyd, dzd;
          const double rs = rsq_image(p, pj[u], S.Lx, S.Ly, S.hLx, S.hLy, dxd, dyd, dzd);
          if (rs > S.R0_2) {
            set_error(S.ctrl, MD_EBOND, i < nn[u] ? i : nn[u], step);
            continue;
          }
          const R gg = -kR / (R02 - (R)rs);
          fx += gg * (R)dxd;
          fy += gg * (R)dyd;
          fz += gg * (R)dzd;
        }
      }
      fb[0][lane] = fx;
      fb[1][lane] = fy;
      fb[2][lane] = fz;
      wzs[lane] = wall_force(S, p);
    }
  } else if (w == 6) {  // the particle's state, for after the barrier
    K = step_constants(S);
    // SR_COUL: the Coulomb force of the step's evaluation is assembled here
    // (as k_coul_store would, the same bits) and held for the next steps
    if (live) {
      pi = S.pos[i];
      const int fl = S.flags[i];
      anchor = fl & 1;
      charged = (fl & 2) != 0;
      if (mode & SR_HELD) {
        if constexpr (sizeof(R) == 4) {
          cx = ((const float *)S.ccx)[i];
          cy = ((const float *)S.ccy)[i];
          cz = ((const float *)S.ccz)[i];
        } else {
          cx = ((const double *)S.ccx)[i];
          cy = ((const double *)S.ccy)[i];
          cz = ((const double *)S.ccz)[i];
        }
      }
      // SR_COUL: the particle's entries load beside its state (one memory
      // round trip; only a charged particle's are defined and used)
      if (coul) cp.load(S, i);
      if (mode & SR_KICK) {
        vx = S.vx[i];
        vy = S.vy[i];
        vz = S.vz[i];
        if (mode & SR_DRIFT) {
          ax0 = S.ax[i];
          ay0 = S.ay[i];
          az0 = S.az[i];
        }
      }
    }
    K.hold();
  } else {  // wave 7
    // the next step's rebuild queue (its list kernels of two steps back are done)
    if ((mode & SR_DRIFT) && blockIdx.x == 0 && lane == 0) *lq_nx = 0ull;
    // SR_BIN: every workgroup takes the decision (the same inputs, the same
    // result) and bins its own particles; the replica's first workgroup
    // queues the rebuild and leaves the bookkeeping to the list kernel, so
    // that no workgroup of this launch writes what the others read
    const bool every = (mode & SR_BIN) != 0;
    if ((mode & SR_DRMAX) && (every || blk == 0)) {  // this step's displacements, all groups
      Ctrl *c = S.ctrl;
      // the accumulated maxima, the rebuild generation and the skins load
      // beside the group maxima, and the group maxima eight per lane at a
      // time: the decision is one memory round trip, shorter than the LJ
      // sums the barrier waits for (no workgroup of this launch writes any
      // of them)
      const double m1 = c->dr_max1, m2 = c->dr_max2;
      const int gen1 = c->gen1;
      const double skin1 = S.skin1, skin2 = S.skin2;
      const int rverlet = S.real_verlet;
      double d1 = 0, d2 = 0;
      const double2 *dp = S.dpart + (size_t)S.par * S.ngrp;
      constexpr int U = 8;
      for (int g0 = 0; g0 < S.ngrp; g0 += 64 * U) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int g = g0 + u * 64 + lane;
          v[u] = g < S.ngrp ? dp[g] : make_double2(0, 0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          d1 = fmax(d1, v[u].x);
          d2 = fmax(d2, v[u].y);
        }
      }
      d1 = wave_max_nonneg(d1);
      d2 = wave_max_nonneg(d2);
      if (every) {
        const double a1 = m1 + d1, a2 = m2 + d2;
        const int f1 = a1 > skin1;
        const int f2 = (a2 > skin2) && rverlet != 0;
        const int buf = (gen1 + 1) & 1;  // the next generation's count buffer
        if (f1) bin_lj_particle(&S, live ? i : -1, buf);
        if (blk == 0 && lane == 0) {
          c->pend_d1 = d1;
          c->pend_d2 = d2;
          c->pend = 1;
          c->flag1 = f1;
          c->flag2 = f2;
          if (f1) lq_push(lq, lq_cap, r, S.NN, buf);
        }
      } else if (lane == 0) {
        c->dr_max1 = m1 + d1;
        c->dr_max2 = m2 + d2;
        int f1 = c->dr_max1 > skin1;
        int f2 = (c->dr_max2 > skin2) && rverlet != 0;
        if (f1) {
          c->dr_max1 = 0;
          c->gen1++;
          lq_push(lq, lq_cap, r, S.NN, c->gen1 & 1);
        }
        if (f2) { c->dr_max2 = 0; c->gen2++; }
        c->flag1 = f1;
        c->flag2 = f2;
      }
    }
  }
  __syncthreads();
  if (w != 6) return;
  double d1 = 0, d2 = 0;
  if (live) {
    const R fx = lj[0][lane] + fb[0][lane], fy = lj[1][lane] + fb[1][lane],
            fz = lj[2][lane] + fb[2][lane];
    double ax = (double)fx, ay = (double)fy, az = (double)fz + wzs[lane];
    if (coul && charged) {  // this step's Coulomb force: held for the next steps and added
      double hx, hy, hz;
      coul_value<R>(S, cp, pi.w, ctot, hx, hy, hz);
      const R rx = (R)hx, ry = (R)hy, rz = (R)hz;
      ((R *)S.ccx)[i] = rx;
      ((R *)S.ccy)[i] = ry;
      ((R *)S.ccz)[i] = rz;
      cx = (double)rx;
      cy = (double)ry;
      cz = (double)rz;
    }
    if (mode & SR_HELD) {
      ax += cx;
      ay += cy;
      az += cz;
    }
    if (mode & (SR_NOISE_ETA | SR_NOISE_PHILOX)) {
      ax += nzs[0][lane] * K.sig;
      ay += nzs[1][lane] * K.sig;
      az += nzs[2][lane] * K.sig;
    }
    if (mode & SR_FIELD) az += pi.w * K.EF;
    if ((mode & SR_ANCHOR) && anchor) ax = ay = az = 0;
    if (!(isfinite(ax) && isfinite(ay) && isfinite(az))) set_error(K.ctrl, MD_ENONFINITE, i, step);
    if (out) {
      out[i] = ax;
      out[(size_t)K.NN + i] = ay;
      out[2 * (size_t)K.NN + i] = az;
    } else {
      K.ax[i] = ax;
      K.ay[i] = ay;
      K.az[i] = az;
      if (mode & SR_KICK) {
#pragma clang fp contract(off)
        if (mode & SR_DRIFT) {  // v holds v(n-1): redo this step's first half kick
          vx = half_kick(K, ax0, vx);
          vy = half_kick(K, ay0, vy);
          vz = half_kick(K, az0, vz);
        }
        const double dt = K.dt;
        vx = (vx + ax * dt / 2.) / K.c1p;
        vy = (vy + ay * dt / 2.) / K.c1p;
        vz = (vz + az * dt / 2.) / K.c1p;
        K.vx[i] = vx;
        K.vy[i] = vy;
        K.vz[i] = vz;
        if (mode & SR_DRIFT) {  // the next step's first half kick and drift, into the other buffer
          double4 p = pi;
          const double dr =
              drift(K, p, half_kick(K, ax, vx), half_kick(K, ay, vy), half_kick(K, az, vz));
          K.pos_nx[i] = p;
          if (K.pos32_nx) K.pos32_nx[i] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
          d1 = dr;
          if (p.w != 0) d2 = dr;
        }
      }
    }
  }
  if (mode & SR_DRIFT) dmax_group(K, K.par ^ 1, blk, d1, d2);
}

// ------------------------------------------------------- Verlet lists

// Forces (clear = 0) or clears (clear = 1) the rebuild flags of every
// replica; forcing also queues every replica's LJ rebuild.  One workgroup.
__global__ void k_set_flags(const Rep *__restrict__ reps, int nrep, int clear,
                            unsigned long long *lq, int cap) {
  if (threadIdx.x == 0) *lq = 0ull;
  __syncthreads();
  for (int r = threadIdx.x; r < nrep; r += blockDim.x) {
    Ctrl *c = reps[r].ctrl;
    if (clear) {
      c->flag1 = c->flag2 = 0;
      continue;
    }
    c->flag1 = 1;
    c->gen1++;
    lq_push(lq, cap, r, reps[r].NN, c->gen1 & 1);
    c->flag2 = reps[r].real_verlet != 0 && reps[r].Nq > 0;
    if (c->flag2) c->gen2++;
  }
}

// Cell of a particle as lj_verlet_list computes it (compt_acc.f90:219-222).
__device__ __forceinline__ bool cell_of(double4 p, const int nc[3], const double rc[3],
                                        double hLx, double hLy, int &a, int &b, int &c) {
  a = (int)((p.x + hLx) / rc[0]);
  b = (int)((p.y + hLy) / rc[1]);
  c = (int)(p.z / rc[2]);
  return a >= 0 && a < nc[0] && b >= 0 && b < nc[1] && c >= 0 && c < nc[2];
}

// Neighbour cell l (0..26) of cell (a,b,c) in the loop order of
// lj_verlet_list (compt_acc.f90:232-255): x outer, y, z inner; x,y periodic,
// z open (-1 when outside).  The reference wraps a neighbour index by one
// period, so with two cells along x (or y) both dp = -1 and dp = +1 name the
// other cell and its members are listed twice, and with one cell all three
// name it (compt_acc.f90:232-248, 545-558): the lists keep those repeats
// (kOnce = false).  The real-space sweep visits every cell once (kOnce).
template <bool kOnce>
__device__ __forceinline__ int stencil_cell(int l, int a, int b, int c, const int nc[3]) {
  const int dp = l / 9 - 1, dq = (l / 3) % 3 - 1;
  if (kOnce) {  // fewer than three cells along a periodic axis: each cell once
    if ((nc[0] == 1 && dp != 0) || (nc[0] == 2 && dp < 0)) return -1;
    if ((nc[1] == 1 && dq != 0) || (nc[1] == 2 && dq < 0)) return -1;
  }
  int p = a + dp, q = b + dq, z = c + l % 3 - 1;
  if (z < 0 || z >= nc[2]) return -1;
  p = p >= nc[0] ? p - nc[0] : (p < 0 ? p + nc[0] : p);
  q = q >= nc[1] ? q - nc[1] : (q < 0 ? q + nc[1] : q);
  return (z * nc[1] + q) * nc[0] + p;
}

// LJ list rebuild (lj_verlet_list, compt_acc.f90:185-314) of the queued
// replicas in three particle-parallel passes: bin, rank within the cell,
// fill.  The kernels stride over the queue's global particle index, so the
// work is proportional to the particles rebuilt, whatever the cell count.

// Binning pass (compt_acc.f90:219-225): cell of each particle, appended to
// the cell's members (order fixed by the rank pass).  Also clears the
// other count buffer at the particle's previous cell, for the next build.

// Particle i of replica *Sp (i < 0: none) into count buffer buf; every lane
// of the wavefront calls it.  Also clears the other buffer at the
// particle's previous cell.
__device__ __forceinline__ void bin_lj_particle(const Rep *Sp, int i, int buf) {
  int *ctr = nullptr;
  int cell = -1;
  if (i >= 0) {
    const Rep &S = *Sp;
    const int ncell = S.nc1[0] * S.nc1[1] * S.nc1[2];
    const int old = S.cell1_of[i];
    if (old >= 0) S.cnt1[(size_t)(buf ^ 1) * ncell + old] = 0;
    int a, b, cz;
    if (cell_of(S.pos[i], S.nc1, S.rcel1, S.hLx, S.hLy, a, b, cz)) {
      cell = (cz * S.nc1[1] + b) * S.nc1[0] + a;
      ctr = S.cnt1 + (size_t)buf * ncell + cell;
    } else {
      set_error(S.ctrl, MD_ECELL, i, -1);
    }
    S.cell1_of[i] = cell;
  }
  const int slot = wave_append(ctr);
  if (ctr) {
    const Rep &S = *Sp;
    if (slot < S.ccap1) S.mem1[(size_t)cell * S.ccap1 + slot] = i;
    else atomicOr(&S.ctrl->ovf, 1);
    // the largest occupancy seen (read first: one address for every particle)
    if (slot >= S.ctrl->max_cell1) atomicMax(&S.ctrl->max_cell1, slot + 1);
  }
}

__device__ __forceinline__ void bin_lj_pass(const Rep *__restrict__ reps, unsigned long long *lq,
                                            int cap) {
  const unsigned long long head = *lq;
  const int ne = min((int)(head >> kQShift), cap);
  const int total = (int)(head & ((1ull << kQShift) - 1));
  const int *qrep = lq_rep(lq), *qst = lq_start(lq, cap);
  for (int base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const int g = base + threadIdx.x;
    const Rep *Sp = reps;
    int i = -1, buf = 0;
    if (g < total) {
      const int e = lq_find(qst, ne, g);
      Sp = reps + lq_entry_rep(qrep[e]);
      buf = lq_entry_buf(qrep[e]);
      i = g - qst[e];
    }
    bin_lj_particle(Sp, i, buf);
  }
}

__global__ __launch_bounds__(kBlock) void k_bin_lj(const Rep *__restrict__ reps,
                                                   unsigned long long *lq, int cap) {
  bin_lj_pass(reps, lq, cap);
}

// Shared 27-cell candidate set of one cell: offs[l] = first candidate of
// stencil cell l (reference order), cids[l] = its index (-1 outside in z).
// Wavefront 0 loads the 27 counts in parallel and scans them.
__device__ __forceinline__ int stencil_offsets(int cell, const int nc[3], const int *scnt, int *offs,
                                               int *cids) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const int a = cell % nc[0], b = (cell / nc[0]) % nc[1], c = cell / (nc[0] * nc[1]);
    int cnt = 0;
    if (lane < 27) {
      const int cl = stencil_cell<true>(lane, a, b, c, nc);
      cids[lane] = cl;
      cnt = cl >= 0 ? scnt[cl] : 0;
    }
    int incl = cnt;
    for (int o = 1; o < 32; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (lane < 27) offs[lane] = incl - cnt;
    if (lane == 26) offs[27] = incl;
  }
  __syncthreads();
  return offs[27];
}

__device__ __forceinline__ int stencil_slot(const int *offs, int t) {
  int l = 0;  // largest l with offs[l] <= t
  for (int k = 16; k > 0; k >>= 1)
    if (l + k < 27 && offs[l + k] <= t) l += k;
  return l;
}

// Fill pass of lj_verlet_list (compt_acc.f90:271-313): 32 lanes per target
// particle.  The candidates are the members of the 27 stencil cells in the
// reference order (stencil order, ascending within a cell): the lanes first
// place each binned member at its rank among its cell's members (members
// are distinct particles) in an LDS stage of kLJStage slots per target (a
// crowded stencil finds the member of each rank by counting instead), then
// test 32 candidates per round and append the hits in that order with a
// ballot, so that lists and every force sum over them do not depend on the
// order of the binning atomics.
// Pairs: i != j and r^2 < rvl^2, r^2 rounded as rij_and_rr rounds it.  The
// list is ELL: slot k of particle i at k*NN + i.  The last kernel of a step
// also advances every replica's step counter.
// One target of the fill pass (below): particle i of replica S (live: a
// target at all), into count generation buf; the 32 lanes of half-wave
// `half` of the wavefront take it.
__device__ __forceinline__ void fill_target(const Rep &S, int i, int buf, bool live, int g,
                                            unsigned long long gmask, unsigned long long below,
                                            int *stage) {
  int a = 0, b = 0, cz = 0;
  double4 pi = make_double4(0, 0, 0, 0);
  if (live) {
    pi = S.pos[i];
    const int cell = S.cell1_of[i];
    live = cell >= 0;
    if (live) {
      a = cell % S.nc1[0];
      b = (cell / S.nc1[0]) % S.nc1[1];
      cz = cell / (S.nc1[0] * S.nc1[1]);
    }
  }
  // stencil cell of lane g (< 27), its count, exclusive offsets in the group
  int cl = -1, cn = 0;
  if (live && g < 27) {
    cl = stencil_cell<false>(g, a, b, cz, S.nc1);
    const int ncell = S.nc1[0] * S.nc1[1] * S.nc1[2];
    if (cl >= 0) cn = min(S.cnt1[(size_t)buf * ncell + cl], S.ccap1);
  }
  int incl = cn;
  for (int o = 1; o < 32; o <<= 1) {
    const int v = __shfl_up(incl, o, 32);
    if (g >= o) incl += v;
  }
  const int excl = incl - cn;
  const int tot = __shfl(incl, 26, 32);
  const int tmax = max(tot, __shfl_xor(tot, 32, 64));
  const bool staged = tot <= kLJStage;
  // stencil slot of candidate t: the largest l with excl_l <= t
  auto slot_of = [&](int t) {
    int l = 0;
    for (int s = 16; s > 0; s >>= 1) {
      const int c = l + s;
      const int ex = __shfl(excl, c < 27 ? c : 26, 32);
      if (c < 27 && ex <= t) l = c;
    }
    return l;
  };
  for (int k0 = 0; k0 < tmax; k0 += 32) {  // binned member -> its place in order
    const int t = k0 + g;
    const int l = slot_of(t);
    const int lcell = __shfl(cl, l, 32), lex = __shfl(excl, l, 32), lcn = __shfl(cn, l, 32);
    if (staged && t < tot) {
      const int *m = S.mem1 + (size_t)lcell * S.ccap1;
      const int j = m[t - lex];
      int rk = 0;
      for (int k = 0; k < lcn; k++) rk += m[k] < j;
      stage[lex + rk] = j;
    }
  }
  wave_sync();
  int n = 0;
  for (int k0 = 0; k0 < tmax; k0 += 32) {
    const int t = k0 + g;
    const int l = slot_of(t);
    const int lcell = __shfl(cl, l, 32), lex = __shfl(excl, l, 32), lcn = __shfl(cn, l, 32);
    bool hit = false;
    int j = 0;
    if (t < tot) {
      if (staged) {
        j = stage[t];
      } else {  // the member of rank t - lex: the one with that many smaller members
        const int *m = S.mem1 + (size_t)lcell * S.ccap1;
        for (int k = 0; k < lcn; k++) {
          const int v = m[k];
          int rk = 0;
          for (int k2 = 0; k2 < lcn; k2++) rk += m[k2] < v;
          if (rk == t - lex) j = v;
        }
      }
      double dx, dy, dz;
      hit = j != i && rsq_image(pi, S.pos[j], S.Lx, S.Ly, S.hLx, S.hLy, dx, dy, dz) < S.rv1sq;
    }
    const unsigned long long mask = __ballot(hit);
    if (hit) {
      const int slot = n + __popcll(mask & below);
      if (slot < S.nbcap1) S.nb1[(size_t)slot * S.NN + i] = j;
    }
    n += __popcll(mask & gmask);
  }
  wave_sync();  // the stage is refilled for the next target
  if (live && g == 0) {
    S.nbn1[i] = n;
    if (n > S.nbcap1) atomicOr(&S.ctrl->ovf, 4);
    if (n > S.ctrl->max_nb1) atomicMax(&S.ctrl->max_nb1, n);
  }
}

__device__ __forceinline__ void build_lj_pass(const Rep *__restrict__ reps, unsigned long long *lq,
                                              int cap, int (*stage_all)[kLJStage]) {
  const unsigned long long head = *lq;
  const int ne = min((int)(head >> kQShift), cap);
  const int total = (int)(head & ((1ull << kQShift) - 1));
  const int *qrep = lq_rep(lq), *qst = lq_start(lq, cap);
  const int lane = threadIdx.x & 63, g = lane & 31, half = lane >> 5;
  const unsigned long long gmask = half ? 0xffffffff00000000ull : 0xffffffffull;
  const unsigned long long below = gmask & ((1ull << lane) - 1ull);
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int *stage = stage_all[threadIdx.x >> 5];
  const int stride = gridDim.x * blockDim.x / 32;
  for (int t0 = 2 * wave; t0 < total; t0 += stride) {
    const int gi = t0 + half;
    const bool in = gi < total;
    int qe = 0, i = 0;
    if (in) {
      const int e = lq_find(qst, ne, gi);
      qe = qrep[e];
      i = gi - qst[e];
    }
    // the wavefront's two targets mostly belong to one queue entry (with one
    // replica, always): then the descriptor is wave-uniform and its fields
    // come by scalar loads instead of a vector load per field on each
    // target's chain of dependent loads (lane 0 holds target t0 < total)
    const int q0 = __builtin_amdgcn_readfirstlane(qe);
    if (__ballot(in && qe != q0) == 0ull)
      fill_target(reps[lq_entry_rep(q0)], i, lq_entry_buf(q0), in, g, gmask, below, stage);
    else
      fill_target(reps[lq_entry_rep(in ? qe : q0)], i, lq_entry_buf(in ? qe : q0), in, g, gmask,
                  below, stage);
  }
}

__global__ __launch_bounds__(kBlock, 8) void k_build_lj(const Rep *__restrict__ reps, int nrep,
                                                     unsigned long long *lq, int cap,
                                                     int inc_step) {
  if (inc_step && blockIdx.x == 0)
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) {
      Ctrl *c = reps[r].ctrl;
      c->step++;
      if (c->pend) {  // the bookkeeping of an SR_BIN step (its flags are set)
        if (c->flag1) {
          c->dr_max1 = 0;
          c->gen1++;
        } else {
          c->dr_max1 = c->dr_max1 + c->pend_d1;
        }
        if (c->flag2) {
          c->dr_max2 = 0;
          c->gen2++;
        } else {
          c->dr_max2 = c->dr_max2 + c->pend_d2;
        }
        c->pend = 0;
      }
    }
  __shared__ int stage_all[kBlock / 32][kLJStage];
  build_lj_pass(reps, lq, cap, stage_all);
}

// Empties a rebuild queue (after a list update outside the step cycle, whose
// next kick would empty it otherwise).
__global__ void k_lq_clear(unsigned long long *lq) { *lq = 0ull; }

// The reference-form real-space list, real_verlet_list
// (compt_acc.f90:491-626): cells of edge >= r_c + skin over the charges,
// binned, sorted, then one thread per charge fills its full list.  Built on
// request (md_build_lists) for every replica that keeps a list, and in the
// step loop at the reference's rebuild steps for the replicas whose
// real-space force walks it (Rep::real_list); the others find real-space
// pairs with the cell sweep of k_real_cells instead.
__global__ __launch_bounds__(kBlock) void k_bin_rl(const Rep *__restrict__ reps, WMap wm) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  Ctrl *c = S.ctrl;
  if (!c->flag2) return;
  const int nblk = wm.blk[r + 1] - wm.blk[r];
  for (int t = (blockIdx.x - wm.blk[r]) * kBlock + threadIdx.x; t < S.Nq; t += nblk * kBlock) {
    const int i = S.charge[t];
    int a, b, cz;
    if (!cell_of(S.pos[i], S.nc2, S.rcel2, S.hLx, S.hLy, a, b, cz)) {
      set_error(c, MD_ECELL, i, -1);
      continue;
    }
    const int cell = (cz * S.nc2[1] + b) * S.nc2[0] + a;
    int *cnt = S.cnt2[c->gen2 & 1];
    const int slot = atomicAdd(&cnt[cell], 1);
    if (slot < S.ccap2) S.mem2[(size_t)cell * S.ccap2 + slot] = t;
    else atomicOr(&c->ovf, 2);
  }
}

// Bitonic sort of one cell's members in LDS (one workgroup per cell).
__device__ void sort_cell_lds(int *m, int n, int *sbuf) {
  int np = 1;
  while (np < n) np <<= 1;
  for (int k = threadIdx.x; k < np; k += blockDim.x) sbuf[k] = k < n ? m[k] : 0x7fffffff;
  __syncthreads();
  for (int size = 2; size <= np; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < np; k += blockDim.x) {
        const int l = k ^ stride;
        if (l > k) {
          const bool up = (k & size) == 0;
          const int x = sbuf[k], y = sbuf[l];
          if ((x > y) == up) {
            sbuf[k] = y;
            sbuf[l] = x;
          }
        }
      }
      __syncthreads();
    }
  for (int k = threadIdx.x; k < n; k += blockDim.x) m[k] = sbuf[k];
}

__global__ __launch_bounds__(kBlock) void k_sort_rl(const Rep *__restrict__ reps, WMap wm) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  Ctrl *c = S.ctrl;
  if (!c->flag2) return;
  const int cell = blockIdx.x - wm.blk[r];
  const int n = min(S.cnt2[c->gen2 & 1][cell], S.ccap2);
  extern __shared__ int sbuf[];
  sort_cell_lds(S.mem2 + (size_t)cell * S.ccap2, n, sbuf);
}

// Fill pass: a wavefront per charge slot t (the replica's wavefronts stride
// over its charges).  The candidates are the members of the 27 stencil cells
// in the reference's loop order, ascending within a cell (k_sort_rl); 64 are
// tested at a time and the hits appended in that order with a ballot.
__global__ __launch_bounds__(kBlock) void k_build_rl(const Rep *__restrict__ reps, WMap wm) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  Ctrl *c = S.ctrl;
  if (!c->flag2) return;
  const int blk = blockIdx.x - wm.blk[r];
  const int nblk = wm.blk[r + 1] - wm.blk[r];
  const int ncell = S.nc2[0] * S.nc2[1] * S.nc2[2];
  const int buf = c->gen2 & 1;
  // clear the other count buffer for the next build
  for (int k = blk * kBlock + threadIdx.x; k < ncell; k += nblk * kBlock) S.cnt2[buf ^ 1][k] = 0;
  const int *cnt = S.cnt2[buf];
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int nwave = nblk * (kBlock / 64);
  for (int t = blk * (kBlock / 64) + (threadIdx.x >> 6); t < S.Nq; t += nwave) {
    const int i = S.charge[t];
    const double4 pi = S.pos[i];
    int a, b, cz;
    if (!cell_of(pi, S.nc2, S.rcel2, S.hLx, S.hLy, a, b, cz)) continue;
    int n = 0;
    for (int l = 0; l < 27; l++) {
      const int cell = stencil_cell<false>(l, a, b, cz, S.nc2);
      if (cell < 0) continue;
      const int cn = min(cnt[cell], S.ccap2);
      const int *cm = S.mem2 + (size_t)cell * S.ccap2;
      for (int u0 = 0; u0 < cn; u0 += 64) {
        const int u = u0 + lane;
        bool hit = false;
        int j = 0;
        if (u < cn) {
          j = S.charge[cm[u]];
          double dx, dy, dz;
          hit = j != i && rsq_image(pi, S.pos[j], S.Lx, S.Ly, S.hLx, S.hLy, dx, dy, dz) < S.rv2sq;
        }
        const unsigned long long mask = __ballot(hit);
        if (hit) {
          const int slot = n + __popcll(mask & below);
          if (slot < S.nbcap2) S.nb2[(size_t)slot * S.Nq + t] = j;
        }
        n += __popcll(mask);
      }
    }
    if (lane == 0) {
      S.nbn2[t] = n;
      if (n > S.nbcap2) atomicOr(&c->ovf, 8);
      if (n > c->max_nb2) atomicMax(&c->max_nb2, n);
    }
  }
}

// ------------------------------------------------------------- SPME

// B-spline weights of splcof (compt_acc.f90:1250-1321).  Returns the mesh
// index of weight 0; w[t], dw[t] for t = 0..p-1 are SP(:,t+1), dSP(:,t+1):
// weight t sits at grid point uuc - p + t.
// O > 0 fixes the order at compile time (all loops unroll); O = 0 takes p_.
template <typename R, int O = 0>
__device__ __forceinline__ int spline(double coord, double kg, int p_, R *w, R *dw) {
  // evaluated as written (no fused multiply-adds): the weights are then the
  // same bits in every kernel that inlines this, whatever the surrounding
  // code lets the compiler contract (the tile and the plane spreading must
  // give the same mesh)
#pragma clang fp contract(off)
  const int p = O > 0 ? O : p_;
  const double uu = kg * coord;
  const double uuc = ceil(uu);
  const R ww = (R)(uu - uuc + 1);
  if (p == 4) {  // closed form, compt_acc.f90:1272-1287
    R s3 = (R)0.5 * ww * ww;
    R s1 = (R)0.5 * ((R)1 - ww) * ((R)1 - ww);
    R s2 = (R)1 - s1 - s3;
    dw[0] = -s1;
    dw[1] = s1 - s2;
    dw[2] = s2 - s3;
    dw[3] = s3;
    w[3] = ww * s3 / (R)3;
    w[2] = ((ww + (R)1) * s2 + ((R)3 - ww) * s3) / (R)3;
    w[0] = ((R)1 - ww) * s1 / (R)3;
    w[1] = (R)1 - w[0] - w[2] - w[3];
  } else {
    // A[t] = M_n(ww + t), t < n, built up from order 2 by the recursion
    // M_n(x) = (x M_{n-1}(x) + (n-x) M_{n-1}(x-1)) / (n-1); the derivative
    // is dM_p(x) = M_{p-1}(x) - M_{p-1}(x-1).
    R A[kMaxOrder];
    A[0] = ww;
    A[1] = (R)1 - ww;
    for (int n = 3; n <= p; n++) {
      if (n == p)
        for (int t = 0; t < p; t++) dw[t] = (t < n - 1 ? A[t] : (R)0) - (t >= 1 ? A[t - 1] : (R)0);
      R B[kMaxOrder];
      for (int t = 0; t < n; t++) {
        const R x = ww + (R)t;                  // argument of M_n for slot t
        const R hi = t < n - 1 ? A[t] : (R)0;   // M_{n-1}(x)
        const R lo = t >= 1 ? A[t - 1] : (R)0;  // M_{n-1}(x-1)
        B[t] = (x * hi + ((R)n - x) * lo) / (R)(n - 1);
      }
      for (int t = 0; t < n; t++) A[t] = B[t];
    }
    // SP(:,j) = M_p(ww + p - j) (compt_acc.f90:1290-1301): weight j-1 is slot p-j
    for (int j = 0; j < p; j++) w[j] = A[p - 1 - j];
    R d[kMaxOrder];
    for (int j = 0; j < p; j++) d[j] = dw[p - 1 - j];
    for (int j = 0; j < p; j++) dw[j] = d[j];
  }
  return (int)uuc - p;
}

__device__ __forceinline__ int wrapk(int v, int K) {
  v %= K;
  return v < 0 ? v + K : v;
}

// First mesh index of a charge's spline weights along one axis, wrapped
// into [0,K): uuc - p of splcof (compt_acc.f90:1269-1270,1304-1308).
__device__ __forceinline__ int spline_base(double coord, double kg, int p, int K) {
  return wrapk((int)ceil(kg * coord) - p, K);
}

// One charge's p^3 products into the LDS tile (int64 fixed point) at
// (x0,y0,z0) with edges E: only the points inside the tile.  The order is a
// template parameter so that every loop unrolls and the weights stay in
// registers.
template <typename R, int O>
__device__ __forceinline__ void tile_add(unsigned long long *acc, double4 p, const double *kg,
                                         const int *K, int x0, int y0, int z0, const int *E) {
  R wx[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder], d[kMaxOrder];
  const int bx = wrapk(spline<R, O>(p.x, kg[0], O, wx, d), K[0]);
  const int by = wrapk(spline<R, O>(p.y, kg[1], O, wy, d), K[1]);
  const int bz = wrapk(spline<R, O>(p.z, kg[2], O, wz, d), K[2]);
  const R q = (R)p.w;
#pragma unroll
  for (int c = 0; c < O; c++) {
    int gz = bz + c;
    gz -= gz >= K[2] ? K[2] : 0;
    const int lz = gz - z0;
    if ((unsigned)lz >= (unsigned)E[2]) continue;
#pragma unroll
    for (int b = 0; b < O; b++) {
      int gy = by + b;
      gy -= gy >= K[1] ? K[1] : 0;
      const int ly = gy - y0;
      if ((unsigned)ly >= (unsigned)E[1]) continue;
      unsigned long long *row = acc + (lz * E[1] + ly) * E[0];
#pragma unroll
      for (int a = 0; a < O; a++) {
        int gx = bx + a;
        gx -= gx >= K[0] ? K[0] : 0;
        const int lx = gx - x0;
        if ((unsigned)lx < (unsigned)E[0]) {
          const double v = (double)(wz[c] * ((wx[a] * q) * wy[b]));
          atomicAdd(row + lx, (unsigned long long)(long long)rint(v * kQScale));
        }
      }
    }
  }
}

// The same with an order per axis (ordr(1:3) unequal), loops not unrolled.
template <typename R>
__device__ __forceinline__ void tile_add_axes(unsigned long long *acc, double4 p, const double *kg,
                                              const int *K, const int *o, int x0, int y0, int z0,
                                              const int *E) {
  R wx[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder], d[kMaxOrder];
  const int bx = wrapk(spline<R, 0>(p.x, kg[0], o[0], wx, d), K[0]);
  const int by = wrapk(spline<R, 0>(p.y, kg[1], o[1], wy, d), K[1]);
  const int bz = wrapk(spline<R, 0>(p.z, kg[2], o[2], wz, d), K[2]);
  const R q = (R)p.w;
  for (int c = 0; c < o[2]; c++) {
    int gz = bz + c;
    gz -= gz >= K[2] ? K[2] : 0;
    const int lz = gz - z0;
    if ((unsigned)lz >= (unsigned)E[2]) continue;
    for (int b = 0; b < o[1]; b++) {
      int gy = by + b;
      gy -= gy >= K[1] ? K[1] : 0;
      const int ly = gy - y0;
      if ((unsigned)ly >= (unsigned)E[1]) continue;
      unsigned long long *row = acc + (lz * E[1] + ly) * E[0];
      for (int a = 0; a < o[0]; a++) {
        int gx = bx + a;
        gx -= gx >= K[0] ? K[0] : 0;
        const int lx = gx - x0;
        if ((unsigned)lx < (unsigned)E[0]) {
          const double v = (double)(wz[c] * ((wx[a] * q) * wy[b]));
          atomicAdd(row + lx, (unsigned long long)(long long)rint(v * kQScale));
        }
      }
    }
  }
}

// MapCharges (compt_acc.f90:1324-1365) by mesh tiles: one workgroup per
// tile accumulates the charges listed for it (those with spline points in
// the tile) in LDS, in 2^-40 fixed point, so the sums do not depend on the
// order of the additions; every mesh point is then written once.
template <typename R>
__device__ __forceinline__ void spread_block(const Rep *__restrict__ reps, WMap wm, int gb,
                                             unsigned long long *acc) {
  const int r = find_rep(wm, gb);
  const Rep &S = reps[r];
  const int tile = gb - wm.blk[r];
  const int o = S.order;
  const int tx = tile % S.ntl[0], ty = (tile / S.ntl[0]) % S.ntl[1], tz = tile / (S.ntl[0] * S.ntl[1]);
  const int E[3] = {tile_edge(S, 0, tx), tile_edge(S, 1, ty), tile_edge(S, 2, tz)};
  const int K[3] = {S.K[0], S.K[1], S.K[2]};
  const int x0 = tile_start(S, 0, tx), y0 = tile_start(S, 1, ty), z0 = tile_start(S, 2, tz);
  const int npt = E[0] * E[1] * E[2];
  for (int k = threadIdx.x; k < npt; k += blockDim.x) acc[k] = 0ull;
  __syncthreads();
  // the tile's sub-lists, concatenated
  int off[kTileSub + 1];
  off[0] = 0;
#pragma unroll
  for (int u = 0; u < kTileSub; u++) off[u + 1] = off[u] + min(S.tcnt[tile * kTileSub + u], S.capt);
  const int n = off[kTileSub];
  const int *mem = S.tmem + (size_t)tile * kTileSub * S.capt;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    int u = 0, ub = 0;
#pragma unroll
    for (int v = 1; v < kTileSub; v++)
      if (c >= off[v]) {
        u = v;
        ub = off[v];
      }
    const double4 p = S.pos[mem[(size_t)u * S.capt + (c - ub)]];
    switch (o) {
      case 3: tile_add<R, 3>(acc, p, S.kg, K, x0, y0, z0, E); break;
      case 4: tile_add<R, 4>(acc, p, S.kg, K, x0, y0, z0, E); break;
      case 5: tile_add<R, 5>(acc, p, S.kg, K, x0, y0, z0, E); break;
      case 6: tile_add<R, 6>(acc, p, S.kg, K, x0, y0, z0, E); break;
      case 7: tile_add<R, 7>(acc, p, S.kg, K, x0, y0, z0, E); break;
      case 8: tile_add<R, 8>(acc, p, S.kg, K, x0, y0, z0, E); break;
      default: tile_add_axes<R>(acc, p, S.kg, K, S.ord, x0, y0, z0, E); break;
    }
  }
  __syncthreads();
  R *q = (R *)S.meshr;
  for (int k = threadIdx.x; k < npt; k += blockDim.x) {
    const int lx = k % E[0], ly = (k / E[0]) % E[1], lz = k / (E[0] * E[1]);
    q[((size_t)(z0 + lz) * K[1] + y0 + ly) * K[0] + x0 + lx] =
        (R)((double)(long long)acc[k] * kQInv);
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_spread(const Rep *__restrict__ reps, WMap wm) {
  __shared__ unsigned long long acc[kTile * kTile * kTile];
  spread_block<R>(reps, wm, blockIdx.x, acc);
}

// LDS FFTs over `nl` lines of length n = 2^ln stored line-major: Stockham
// autosort passes (natural order in and out), radix 8 while three or more
// factors of two remain, then radix 4 or 2, each butterfly held in
// registers.  A pass reads one buffer and writes the other; fft_lines
// returns the buffer that holds the result.  tw[k] = exp(-2 pi i k / n),
// k < n; SGN = -1 forward, +1 inverse (unnormalised, FFTW's convention).
template <typename C, int SGN>
__device__ __forceinline__ C rot90(C a) {  // a * (SGN i)
  C r;
  if (SGN < 0) { r.x = a.y; r.y = -a.x; }
  else { r.x = -a.y; r.y = a.x; }
  return r;
}

template <typename C, int SGN>
__device__ __forceinline__ C mul_w8(C a) {  // a * exp(SGN i pi / 4)
  using R = decltype(a.x);
  const R h = (R)0.70710678118654752440;
  C r;
  if (SGN < 0) { r.x = (a.x + a.y) * h; r.y = (a.y - a.x) * h; }
  else { r.x = (a.x - a.y) * h; r.y = (a.x + a.y) * h; }
  return r;
}

template <typename C, int SGN>
__device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3) {
  const C t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = rot90<C, SGN>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename C, int SGN, int R>
__device__ __forceinline__ void dft_reg(C *v) {
  using T = decltype(v[0].x);
  if constexpr (R == 3) {  // X1,2 = v0 - (v1 + v2)/2 +- (SGN i) sqrt(3)/2 (v1 - v2)
    const C s = cadd(v[1], v[2]);
    C m = v[0], d = rot90<C, SGN>(csub(v[1], v[2]));
    m.x -= (T)0.5 * s.x;
    m.y -= (T)0.5 * s.y;
    const T c = (T)0.86602540378443864676;
    d.x *= c;
    d.y *= c;
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, d);
    v[2] = csub(m, d);
  } else if constexpr (R == 5) {
    // with w = exp(SGN 2 pi i / 5): X1,4 = v0 + c1 s1 + c2 s2 +- (SGN i)(n1 d1 + n2 d2),
    // X2,3 = v0 + c2 s1 + c1 s2 +- (SGN i)(n2 d1 - n1 d2); s1,d1 = v1 +- v4, s2,d2 = v2 +- v3
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T n1 = (T)0.95105651629515357212, n2 = (T)0.58778525229247312917;
    const C s1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]), s2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    C a1, a2, b1, b2;
    a1.x = v[0].x + c1 * s1.x + c2 * s2.x;
    a1.y = v[0].y + c1 * s1.y + c2 * s2.y;
    a2.x = v[0].x + c2 * s1.x + c1 * s2.x;
    a2.y = v[0].y + c2 * s1.y + c1 * s2.y;
    b1.x = n1 * d1.x + n2 * d2.x;
    b1.y = n1 * d1.y + n2 * d2.y;
    b2.x = n2 * d1.x - n1 * d2.x;
    b2.y = n2 * d1.y - n1 * d2.y;
    b1 = rot90<C, SGN>(b1);
    b2 = rot90<C, SGN>(b2);
    v[0] = cadd(v[0], cadd(s1, s2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  } else if constexpr (R == 2) {
    const C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  } else if constexpr (R == 4) {
    dft4<C, SGN>(v[0], v[1], v[2], v[3]);
  } else {  // 8: two DFT4 over the even and odd points, then W8 butterflies
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<C, SGN>(e0, e1, e2, e3);
    dft4<C, SGN>(o0, o1, o2, o3);
    o1 = mul_w8<C, SGN>(o1);
    o2 = rot90<C, SGN>(o2);
    o3 = rot90<C, SGN>(mul_w8<C, SGN>(o3));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
  }
}

// One pass: blocks of Ns outputs already transformed; radix R joins R of
// them (input stride n/R, output stride Ns) after the twiddles of the block.
template <typename C, int SGN, int R>
__device__ __forceinline__ void stockham_pass(const C *src, C *dst, int ln, int nl, int Ns,
                                              const C *__restrict__ tw, int ls) {
  const int n = 1 << ln, nr = n / R;
  const int lnr = ln - (R == 8 ? 3 : R == 4 ? 2 : 1);
  for (int t = threadIdx.x; t < nl * nr; t += blockDim.x) {
    const int line = t >> lnr, j = t & (nr - 1);
    const int k = j & (Ns - 1);
    const C *s = src + line * ls + j;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = s[r * nr];
    if (Ns > 1) {
      const int step = k * (n / (Ns * R));
#pragma unroll
      for (int r = 1; r < R; r++) {
        C w = tw[r * step];
        if (SGN > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    dft_reg<C, SGN, R>(v);
    C *d = dst + line * ls + (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) d[r * Ns] = v[r];
  }
}

template <typename C, int SGN>
__device__ C *fft_lines(C *a, C *b, int ln, int nl, const C *__restrict__ tw, int ls = 0) {
  if (ls == 0) ls = 1 << ln;  // line start stride (points)
  int rem = ln, Ns = 1;
  C *src = a, *dst = b;
  while (rem > 0) {
    if (rem >= 3) {
      stockham_pass<C, SGN, 8>(src, dst, ln, nl, Ns, tw, ls);
      Ns <<= 3;
      rem -= 3;
    } else if (rem == 2) {
      stockham_pass<C, SGN, 4>(src, dst, ln, nl, Ns, tw, ls);
      Ns <<= 2;
      rem -= 2;
    } else {
      stockham_pass<C, SGN, 2>(src, dst, ln, nl, Ns, tw, ls);
      Ns <<= 1;
      rem -= 1;
    }
    __syncthreads();
    C *x = src;
    src = dst;
    dst = x;
  }
  return src;
}

// Lines of any length n = fac[0] fac[1] ... : one Stockham pass per factor
// (the same autosort indexing with divisions in place of shifts), radix 2,
// 3, 4, 5 and 8 as butterflies in registers; any other factor (a prime) by
// a pass that gives each thread one output, a DFT of the factor's inputs.
template <typename C, int SGN, int R>
__device__ __forceinline__ void stockham_pass_n(const C *src, C *dst, int n, int nl, int Ns,
                                                const C *__restrict__ tw, int ls) {
  const int nr = n / R, q = n / (Ns * R);
  for (int t = threadIdx.x; t < nl * nr; t += blockDim.x) {
    const int line = t / nr, j = t - line * nr;
    const int k = j % Ns;
    const C *s = src + line * ls + j;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = s[r * nr];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        C w = tw[r * k * q];
        if (SGN > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    dft_reg<C, SGN, R>(v);
    C *d = dst + line * ls + (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) d[r * Ns] = v[r];
  }
}

template <typename C, int SGN>
__device__ __forceinline__ void dft_pass_n(const C *src, C *dst, int n, int nl, int Ns,
                                           const C *__restrict__ tw, int ls, int R) {
  const int nr = n / R, q = n / (Ns * R);
  for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
    const int line = t / n, o = t - line * n;
    const int sidx = o / nr, j = o - sidx * nr;  // output sidx of butterfly j
    const int k = j % Ns;
    const int step = (k + sidx * Ns) * q;  // < n: the twiddle and DFT exponent of input 1
    const C *sl = src + line * ls + j;
    C acc;
    acc.x = 0;
    acc.y = 0;
    int idx = 0;
    for (int r = 0; r < R; r++) {
      C w = tw[idx];
      if (SGN > 0) w.y = -w.y;
      acc = cadd(acc, cmul(sl[r * nr], w));
      idx += step;
      idx -= idx >= n ? n : 0;
    }
    dst[line * ls + (j - k) * R + k + sidx * Ns] = acc;
  }
}

template <typename C, int SGN>
__device__ C *fft_lines_mixed(C *a, C *b, int n, int nl, const C *__restrict__ tw, int ls,
                              const int *fac, int nfac) {
  if (ls == 0) ls = n;
  int Ns = 1;
  C *src = a, *dst = b;
  for (int f = 0; f < nfac; f++) {
    const int R = fac[f];
    switch (R) {
      case 2: stockham_pass_n<C, SGN, 2>(src, dst, n, nl, Ns, tw, ls); break;
      case 3: stockham_pass_n<C, SGN, 3>(src, dst, n, nl, Ns, tw, ls); break;
      case 4: stockham_pass_n<C, SGN, 4>(src, dst, n, nl, Ns, tw, ls); break;
      case 5: stockham_pass_n<C, SGN, 5>(src, dst, n, nl, Ns, tw, ls); break;
      case 8: stockham_pass_n<C, SGN, 8>(src, dst, n, nl, Ns, tw, ls); break;
      default: dft_pass_n<C, SGN>(src, dst, n, nl, Ns, tw, ls, R); break;
    }
    Ns *= R;
    __syncthreads();
    C *x = src;
    src = dst;
    dst = x;
  }
  return src;
}

// The lines of mesh axis d: power-of-two lengths take fft_lines.  tw is
// the axis's twiddle table staged in LDS (stage_twiddles).
template <typename C, int SGN>
__device__ __forceinline__ C *fft_axis(const Rep &S, int d, const C *tw, C *a, C *b, int nl, int ls = 0) {
  if (S.pw2[d]) return fft_lines<C, SGN>(a, b, S.lK[d], nl, tw, ls);
  return fft_lines_mixed<C, SGN>(a, b, S.K[d], nl, tw, ls, S.fac[d], S.nfac[d]);
}

// Copies axis d's twiddle table (K_d complex) into LDS; the kernel's barrier
// after its data loads covers it.
template <typename C>
__device__ __forceinline__ void stage_twiddles(const Rep &S, int d, C *dst) {
  const C *tw = (const C *)S.tw[d];
  for (int k = threadIdx.x; k < S.K[d]; k += blockDim.x) dst[k] = tw[k];
}

// Stores st(t, ld(t)) for t = threadIdx.x, threadIdx.x + blockDim.x, ...
// below n, in batches of U elements per thread whose loads are all issued
// before the batch's first store: a global-to-LDS copy then waits one memory
// round trip per batch, not one per element (a plain loop waits for each
// load before its store).
template <int U, typename Ld, typename St>
__device__ __forceinline__ void copy_batched(int n, Ld ld, St st) {
  using T = decltype(ld(0));
  for (int t0 = threadIdx.x; t0 < n; t0 += U * (int)blockDim.x) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = t0 + u * (int)blockDim.x;
      if (t < n) v[u] = ld(t);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = t0 + u * (int)blockDim.x;
      if (t < n) st(t, v[u]);
    }
  }
}

// The x and y (or one axis') twiddle tables of a plane or line kernel,
// loaded before the kernel's own copy so that both wait in the same memory
// round trip: one element per thread where the table fits the workgroup,
// else staged by stage_twiddles.  stage() stores what load() fetched.
template <typename C>
struct TwPair {
  C a, b;
  bool one;
  __device__ __forceinline__ void load(const Rep &S, int da, int db) {
    const int k = threadIdx.x;
    one = S.K[da] <= (int)blockDim.x && (db < 0 || S.K[db] <= (int)blockDim.x);
    a = b = C{};
    if (one && k < S.K[da]) a = ((const C *)S.tw[da])[k];
    if (one && db >= 0 && k < S.K[db]) b = ((const C *)S.tw[db])[k];
  }
  __device__ __forceinline__ void stage(const Rep &S, int da, C *dst_a, int db, C *dst_b) const {
    const int k = threadIdx.x;
    if (one) {
      if (k < S.K[da]) dst_a[k] = a;
      if (db >= 0 && k < S.K[db]) dst_b[k] = b;
    } else {
      stage_twiddles(S, da, dst_a);
      if (db >= 0) stage_twiddles(S, db, dst_b);
    }
  }
};

// Forward x pass: pairs of real mesh rows as one complex FFT of length K1,
// split into two half spectra.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_fftx_fwd(const Rep *__restrict__ reps, WMap wm) {
  using C = typename CT<R>::C;
  extern __shared__ char smem[];
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int K1 = S.K[0], Kh = S.Kh;
  const int nrow = S.K[1] * S.K[2], npairs = (nrow + 1) / 2;  // an odd count leaves a row alone
  const int P = S.px_tile;
  C *ba = (C *)smem, *bb = ba + (size_t)P * K1, *twx = bb + (size_t)P * K1;
  const int pair0 = (blockIdx.x - wm.blk[r]) * P;
  const int nl = min(P, npairs - pair0);
  const R *q = (const R *)S.meshr;
  stage_twiddles(S, 0, twx);
  for (int t = threadIdx.x; t < nl * K1; t += blockDim.x) {
    const int l = t / K1, x = t - l * K1;
    const size_t row = 2 * (size_t)(pair0 + l);
    C v;
    v.x = q[row * K1 + x];
    v.y = row + 1 < (size_t)nrow ? q[(row + 1) * K1 + x] : (R)0;
    ba[l * K1 + x] = v;
  }
  __syncthreads();
  const C *buf = fft_axis<C, -1>(S, 0, twx, ba, bb, nl);
  C *spec = (C *)S.spec;
  for (int t = threadIdx.x; t < nl * Kh; t += blockDim.x) {
    const int l = t / Kh, k = t - l * Kh;
    const C zk = buf[l * K1 + k], zm = buf[l * K1 + (k == 0 ? 0 : K1 - k)];
    C A, B;  // A = (Z_k + conj Z_-k)/2, B = (Z_k - conj Z_-k)/(2i)
    A.x = (R)0.5 * (zk.x + zm.x);
    A.y = (R)0.5 * (zk.y - zm.y);
    B.x = (R)0.5 * (zk.y + zm.y);
    B.y = (R)0.5 * (zm.x - zk.x);
    const size_t row = 2 * (size_t)(pair0 + l);
    spec[row * Kh + k] = A;
    if (row + 1 < (size_t)nrow) spec[(row + 1) * Kh + k] = B;
  }
}

// Inverse x pass: two half spectra -> one complex inverse FFT -> two real
// rows of the potential mesh.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_fftx_inv(const Rep *__restrict__ reps, WMap wm) {
  using C = typename CT<R>::C;
  extern __shared__ char smem[];
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int K1 = S.K[0], Kh = S.Kh;
  const int nrow = S.K[1] * S.K[2], npairs = (nrow + 1) / 2;
  const int P = S.px_tile;
  C *ba = (C *)smem, *bb = ba + (size_t)P * K1, *twx = bb + (size_t)P * K1;
  const int pair0 = (blockIdx.x - wm.blk[r]) * P;
  const int nl = min(P, npairs - pair0);
  const C *spec = (const C *)S.spec;
  stage_twiddles(S, 0, twx);
  for (int t = threadIdx.x; t < nl * K1; t += blockDim.x) {
    const int l = t / K1, k = t - l * K1;
    const size_t row = 2 * (size_t)(pair0 + l);
    const bool two = row + 1 < (size_t)nrow;
    C A, B, z;
    B.x = B.y = (R)0;
    if (k < Kh) {
      A = spec[row * Kh + k];
      if (two) B = spec[(row + 1) * Kh + k];
    } else {
      A = spec[row * Kh + (K1 - k)];
      if (two) B = spec[(row + 1) * Kh + (K1 - k)];
      A.y = -A.y;
      B.y = -B.y;
    }
    z.x = A.x - B.y;  // Z = A + iB
    z.y = A.y + B.x;
    ba[l * K1 + k] = z;
  }
  __syncthreads();
  const C *buf = fft_axis<C, +1>(S, 0, twx, ba, bb, nl);
  R *u = (R *)S.meshu;
  for (int t = threadIdx.x; t < nl * K1; t += blockDim.x) {
    const int l = t / K1, x = t - l * K1;
    const C v = buf[l * K1 + x];
    const size_t row = 2 * (size_t)(pair0 + l);
    u[row * K1 + x] = v.x;
    if (row + 1 < (size_t)nrow) u[(row + 1) * K1 + x] = v.y;
  }
}

// y pass (forward or inverse), on the half spectrum: tiles of Ty adjacent
// kx lines for one kz.
template <typename R, int SGN>
__global__ __launch_bounds__(kBlock) void k_ffty(const Rep *__restrict__ reps, WMap wm) {
  using C = typename CT<R>::C;
  extern __shared__ char smem[];
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int K2 = S.K[1], Kh = S.Kh, T = S.py_tile, KP = plane_line_stride(K2);
  C *ba = (C *)smem, *bb = ba + (size_t)T * KP, *twy = bb + (size_t)T * KP;
  stage_twiddles(S, 1, twy);
  const int tiles = (Kh + T - 1) / T;
  const int b = blockIdx.x - wm.blk[r];
  const int z = b / tiles, x0 = (b - z * tiles) * T;
  const int nl = min(T, Kh - x0);
  C *spec = (C *)S.spec + (size_t)z * K2 * Kh + x0;
  for (int t = threadIdx.x; t < nl * K2; t += blockDim.x) {
    const int y = t / nl, l = t - y * nl;
    ba[l * KP + y] = spec[(size_t)y * Kh + l];
  }
  __syncthreads();
  const C *buf = fft_axis<C, SGN>(S, 1, twy, ba, bb, nl, KP);
  for (int t = threadIdx.x; t < nl * K2; t += blockDim.x) {
    const int y = t / nl, l = t - y * nl;
    spec[(size_t)y * Kh + l] = buf[l * KP + y];
  }
}

// Forward x and y passes of one z-plane in LDS: pairs of real rows -> one
// complex FFT of length K1 each, split into two half spectra stored
// line-major by kx, then the y FFT of each kx line.  Two LDS regions of
// plane_region points: the passes alternate between them.
// The forward x and y passes of z-plane z whose rows are in ra (pairs of
// rows as complex lines), with the twiddles staged: the spectrum plane.
template <typename R, typename C>
__device__ __forceinline__ void plane_fwd_lds(const Rep &S, int z, C *ra, C *rb, const C *twx,
                                              const C *twy) {
  const int K1 = S.K[0], K2 = S.K[1], Kh = S.Kh, H = (K2 + 1) / 2;
  const int KP = plane_line_stride(K2);
  const C *A = fft_axis<C, -1>(S, 0, twx, ra, rb, H);
  C *P = A == ra ? rb : ra;
  for (int t = threadIdx.x; t < H * Kh; t += blockDim.x) {
    const int l = t / Kh, k = t - l * Kh;
    const C zk = A[l * K1 + k], zm = A[l * K1 + (k == 0 ? 0 : K1 - k)];
    C a, b;
    a.x = (R)0.5 * (zk.x + zm.x);
    a.y = (R)0.5 * (zk.y - zm.y);
    b.x = (R)0.5 * (zk.y + zm.y);
    b.y = (R)0.5 * (zm.x - zk.x);
    P[k * KP + 2 * l] = a;
    if (2 * l + 1 < K2) P[k * KP + 2 * l + 1] = b;
  }
  __syncthreads();
  const C *Y = fft_axis<C, -1>(S, 1, twy, P, (C *)A, Kh, KP);
  C *spec = (C *)S.spec + (size_t)z * K2 * Kh;
  for (int t = threadIdx.x; t < K2 * Kh; t += blockDim.x) {
    const int y = t / Kh, k = t - y * Kh;
    spec[t] = Y[k * KP + y];
  }
}

template <typename R>
__global__ __launch_bounds__(kPlaneBlock) void k_plane_fwd(const Rep *__restrict__ reps, WMap wm) {
  using C = typename CT<R>::C;
  extern __shared__ char smem[];
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int z = blockIdx.x - wm.blk[r];
  const int K1 = S.K[0], K2 = S.K[1], Kh = S.Kh, H = (K2 + 1) / 2;  // row pairs (odd K2: one alone)
  C *ra = (C *)smem, *rb = ra + plane_region(K1, K2, Kh);
  C *twx = rb + plane_region(K1, K2, Kh), *twy = twx + K1;
  TwPair<C> tw;
  tw.load(S, 0, 1);
  const R *q = (const R *)S.meshr + (size_t)z * K2 * K1;
  copy_batched<4>(
      H * K1,
      [&](int t) {
        const int l = t / K1, x = t - l * K1;
        C v;
        v.x = q[(size_t)(2 * l) * K1 + x];
        v.y = 2 * l + 1 < K2 ? q[(size_t)(2 * l + 1) * K1 + x] : (R)0;
        return v;
      },
      [&](int t, C v) { ra[t] = v; });
  tw.stage(S, 0, twx, 1, twy);
  __syncthreads();
  plane_fwd_lds<R>(S, z, ra, rb, twx, twy);
}

// One charge's products in z-plane z (weight c = z - its first z index)
// into the plane's LDS accumulator (int64 fixed point, K1 x K2, x fastest):
// the same products, in the same arithmetic, as tile_add's for that plane.
template <typename R, int O>
__device__ __forceinline__ void plane_add(unsigned long long *acc, double4 p, const double *kg,
                                          const int *K, int z) {
  R wx[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder], d[kMaxOrder];
  const int bx = wrapk(spline<R, O>(p.x, kg[0], O, wx, d), K[0]);
  const int by = wrapk(spline<R, O>(p.y, kg[1], O, wy, d), K[1]);
  const int bz = wrapk(spline<R, O>(p.z, kg[2], O, wz, d), K[2]);
  const int c = wrapk(z - bz, K[2]);
  if (c >= O) return;
  const R q = (R)p.w;
#pragma unroll
  for (int b = 0; b < O; b++) {
    int gy = by + b;
    gy -= gy >= K[1] ? K[1] : 0;
    unsigned long long *row = acc + gy * K[0];
#pragma unroll
    for (int a = 0; a < O; a++) {
      int gx = bx + a;
      gx -= gx >= K[0] ? K[0] : 0;
      const double v = (double)(wz[c] * ((wx[a] * q) * wy[b]));
      atomicAdd(row + gx, (unsigned long long)(long long)rint(v * kQScale));
    }
  }
}

// The same with an order per axis (as tile_add_axes).
template <typename R>
__device__ __forceinline__ void plane_add_axes(unsigned long long *acc, double4 p, const double *kg,
                                               const int *K, const int *o, int z) {
  R wx[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder], d[kMaxOrder];
  const int bx = wrapk(spline<R, 0>(p.x, kg[0], o[0], wx, d), K[0]);
  const int by = wrapk(spline<R, 0>(p.y, kg[1], o[1], wy, d), K[1]);
  const int bz = wrapk(spline<R, 0>(p.z, kg[2], o[2], wz, d), K[2]);
  const int c = wrapk(z - bz, K[2]);
  if (c >= o[2]) return;
  const R q = (R)p.w;
  for (int b = 0; b < o[1]; b++) {
    int gy = by + b;
    gy -= gy >= K[1] ? K[1] : 0;
    unsigned long long *row = acc + gy * K[0];
    for (int a = 0; a < o[0]; a++) {
      int gx = bx + a;
      gx -= gx >= K[0] ? K[0] : 0;
      const double v = (double)(wz[c] * ((wx[a] * q) * wy[b]));
      atomicAdd(row + gx, (unsigned long long)(long long)rint(v * kQScale));
    }
  }
}

// MapCharges (compt_acc.f90:1324-1365) fused with the forward x and y
// passes, one workgroup per z-plane: the charges whose first z index is
// one of the order's planes up to z (k_bin_q's plane lists) add their
// products in this plane to an LDS accumulator in 2^-40 fixed point, which
// then becomes the plane's rows, exactly the values the tile spreading
// writes, so the spectrum is bitwise that of k_spread + k_plane_fwd.  The
// accumulator shares the two FFT regions (it is read out into registers
// before the rows are written); the charge mesh is also stored, for the
// mesh energy.
template <typename R>
__device__ __forceinline__ void plane_spread_block(const Rep *__restrict__ reps, WMap wm, int gb,
                                                   char *smem, int store_q) {
  using C = typename CT<R>::C;
  const int r = find_rep(wm, gb);
  const Rep &S = reps[r];
  const int z = gb - wm.blk[r];
  const int K1 = S.K[0], K2 = S.K[1], Kh = S.Kh;
  C *ra = (C *)smem, *rb = ra + plane_region(K1, K2, Kh);
  C *twx = rb + plane_region(K1, K2, Kh), *twy = twx + K1;
  unsigned long long *acc = (unsigned long long *)smem;  // K1 K2 words within ra and rb
  const int npt = K1 * K2;
  TwPair<C> tw;
  tw.load(S, 0, 1);
  for (int k = threadIdx.x; k < npt; k += blockDim.x) acc[k] = 0ull;
  // the lists of first z indices z - c, c = 0..order-1, each of kTileSub
  // sub-lists: sub-list j = c kTileSub + u, its exclusive offset off[j]
  // (wavefront 0: a lane per sub-list, then a scan)
  // (the table follows the twiddles in the dynamic LDS: the kernel has no
  // static LDS, which would shift the dynamic base off the 8-byte alignment
  // of the accumulator's 64-bit atomics)
  static_assert(kMaxOrder * kTileSub <= 64, "a lane per sub-list");
  static_assert((kMaxOrder * kTileSub + 1) * sizeof(int) <= kPlaneSpreadTable, "the table's LDS");
  int *off = (int *)(twy + K2);
  const int oz = S.ord[2], K[3] = {S.K[0], S.K[1], S.K[2]}, nsub = oz * kTileSub;
  if (threadIdx.x < 64) {
    const int j = threadIdx.x;
    const int c = j / kTileSub, u = j - c * kTileSub;
    const int cnt = j < nsub ? min(S.tcnt[wrapk(z - c, S.K[2]) * kTileSub + u], S.capt) : 0;
    int incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (j >= o) incl += v;
    }
    if (j < nsub) off[j] = incl - cnt;
    if (j == nsub - 1) off[nsub] = incl;
  }
  tw.stage(S, 0, twx, 1, twy);
  __syncthreads();  // the accumulator is clear, the table and twiddles staged
  const int n = off[nsub];
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    int j = 0;  // the sub-list holding entry t: the largest j with off[j] <= t
    for (int st = 32; st > 0; st >>= 1)
      if (j + st < nsub && off[j + st] <= t) j += st;
    const int c = j / kTileSub, u = j - c * kTileSub, lz = wrapk(z - c, S.K[2]);
    const double4 p = S.pos[S.tmem[((size_t)lz * kTileSub + u) * S.capt + (t - off[j])]];
    switch (S.order) {
      case 3: plane_add<R, 3>(acc, p, S.kg, K, z); break;
      case 4: plane_add<R, 4>(acc, p, S.kg, K, z); break;
      case 5: plane_add<R, 5>(acc, p, S.kg, K, z); break;
      case 6: plane_add<R, 6>(acc, p, S.kg, K, z); break;
      case 7: plane_add<R, 7>(acc, p, S.kg, K, z); break;
      case 8: plane_add<R, 8>(acc, p, S.kg, K, z); break;
      default: plane_add_axes<R>(acc, p, S.kg, K, S.ord, z); break;
    }
  }
  __syncthreads();
  // read the accumulator out (every word before any row is written: they
  // share the memory), store the charge mesh plane, then the rows as pairs
  R val[kPlaneSpreadPer];
#pragma unroll
  for (int k = 0; k < kPlaneSpreadPer; k++) {
    const int t = threadIdx.x + k * kPlaneBlock;
    val[k] = t < npt ? (R)((double)(long long)acc[t] * kQInv) : (R)0;
  }
  __syncthreads();
  R *qm = (R *)S.meshr + (size_t)z * K2 * K1;
#pragma unroll
  for (int k = 0; k < kPlaneSpreadPer; k++) {
    const int t = threadIdx.x + k * kPlaneBlock;
    if (t < npt) {
      if (store_q) qm[t] = val[k];
      const int y = t / K1, x = t - y * K1;
      R *row = (R *)&ra[(y >> 1) * K1 + x];
      row[y & 1] = val[k];
      if ((K2 & 1) && y == K2 - 1) row[1] = (R)0;  // a lone last row
    }
  }
  __syncthreads();
  plane_fwd_lds<R>(S, z, ra, rb, twx, twy);
}

template <typename R>
__global__ __launch_bounds__(kPlaneBlock) void k_plane_spread(const Rep *__restrict__ reps, WMap wm,
                                                              int store_q) {
  extern __shared__ char smem[];
  plane_spread_block<R>(reps, wm, blockIdx.x, smem, store_q);
}

// Inverse y and x passes of one z-plane: inverse y FFT of each kx line,
// pairs of rows Z = A + iB rebuilt over the full x range (A, B Hermitian),
// inverse x FFT, real and imaginary parts are the two rows.
template <typename R>
__global__ __launch_bounds__(kPlaneBlock) void k_plane_inv(const Rep *__restrict__ reps, WMap wm) {
  using C = typename CT<R>::C;
  extern __shared__ char smem[];
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int z = blockIdx.x - wm.blk[r];
  const int K1 = S.K[0], K2 = S.K[1], Kh = S.Kh, H = (K2 + 1) / 2;
  const int KP = plane_line_stride(K2);
  C *ra = (C *)smem, *rb = ra + plane_region(K1, K2, Kh);
  C *twx = rb + plane_region(K1, K2, Kh), *twy = twx + K1;
  TwPair<C> tw;
  tw.load(S, 0, 1);
  const C *spec = (const C *)S.spec + (size_t)z * K2 * Kh;
  copy_batched<8>(
      K2 * Kh, [&](int t) { return spec[t]; },
      [&](int t, C v) {
        const int y = t / Kh, k = t - y * Kh;
        ra[k * KP + y] = v;
      });
  tw.stage(S, 0, twx, 1, twy);
  __syncthreads();
  const C *P = fft_axis<C, +1>(S, 1, twy, ra, rb, Kh, KP);
  C *A = P == ra ? rb : ra;
  for (int t = threadIdx.x; t < H * K1; t += blockDim.x) {
    const int l = t / K1, k = t - l * K1;
    const bool two = 2 * l + 1 < K2;
    C a, b, v;
    b.x = b.y = (R)0;
    if (k < Kh) {
      a = P[k * KP + 2 * l];
      if (two) b = P[k * KP + 2 * l + 1];
    } else {
      a = P[(K1 - k) * KP + 2 * l];
      if (two) b = P[(K1 - k) * KP + 2 * l + 1];
      a.y = -a.y;
      b.y = -b.y;
    }
    v.x = a.x - b.y;
    v.y = a.y + b.x;
    A[l * K1 + k] = v;
  }
  __syncthreads();
  const C *X = fft_axis<C, +1>(S, 0, twx, A, (C *)P, H);
  R *u = (R *)S.meshu + (size_t)z * K2 * K1;
  for (int t = threadIdx.x; t < H * K1; t += blockDim.x) {
    const int l = t / K1, x = t - l * K1;
    const C v = X[l * K1 + x];
    u[(size_t)(2 * l) * K1 + x] = v.x;
    if (2 * l + 1 < K2) u[(size_t)(2 * l + 1) * K1 + x] = v.y;
  }
}

// z pass with the convolution of pmeOrthoConvBC (compt_acc.f90:1382-1396):
// forward FFT along z, multiply by BC, inverse FFT along z; tiles of Tz
// adjacent (ky, kx) lines, contiguous in the half-spectrum layout.
// Die-grouped order of n work units: the workgroups of a launch are handed
// to the kXCD dies in turn (workgroup b to die b % kXCD), and each die has
// its own L2.  Local workgroups j, j + kXCD, ..., j + (G - 1) kXCD, which
// share a die, take G consecutive units, so that units reading neighbouring
// pieces of one cache line read it through one L2.  A permutation of
// [0, n): whole runs of G kXCD workgroups are regrouped, the rest keep
// their order.
template <int G>
__device__ __forceinline__ int die_grouped(int j, int n) {
  constexpr int run = G * kXCD;
  if (j >= n / run * run) return j;
  const int s = j / kXCD, x = j % kXCD;
  return (s / G) * run + x * G + s % G;
}

template <typename R>
__device__ __forceinline__ void fftz_block(const Rep *__restrict__ reps, WMap wm, int gb, char *smem) {
  using C = typename CT<R>::C;
  const int r = find_rep(wm, gb);
  const Rep &S = reps[r];
  const int K3 = S.K[2], T = S.pz_tile, KS = zline_stride(K3);
  C *ba = (C *)smem, *bb = ba + (size_t)T * KS, *twz = bb + (size_t)T * KS;
  TwPair<C> tw;
  tw.load(S, 2, -1);
  // four line groups per die: 32 lines of 8-byte complex values (FP32
  // path, 8 lines a group) are two whole 128-byte pieces of each z-row of
  // the spectrum and one of the influence function
  const int f0 = die_grouped<4>(gb - wm.blk[r], wm.blk[r + 1] - wm.blk[r]) * T;
  const size_t zs = (size_t)S.K[1] * S.Kh;
  const int nl = min(T, (int)zs - f0);
  C *spec = (C *)S.spec + f0;
  const R *bc = (const R *)S.bc + f0;
  // Thread-fixed lines when the group is whole and the workgroup holds a
  // whole number of its rows (the usual case): a thread keeps line
  // l = tid % nl and steps through z by blockDim / nl, with no division per
  // element; its influence-function values load with the spectrum and stay
  // in registers until the multiply (kZPer per thread at most).  Otherwise
  // element t of the tile is (z, l) = (t / nl, t % nl).
  constexpr int kZPer = 8;
  const int zstep = (int)blockDim.x / nl;
  const bool fixed = nl == T && zstep * nl == (int)blockDim.x && K3 <= kZPer * zstep;
  R g[kZPer];
  if (fixed) {
    const int l = threadIdx.x % nl, z0 = threadIdx.x / nl;
    C v[kZPer];
#pragma unroll
    for (int u = 0; u < kZPer; u++) {
      const int z = z0 + u * zstep;
      if (z < K3) {
        v[u] = spec[(size_t)z * zs + l];
        g[u] = bc[(size_t)z * zs + l];
      }
    }
#pragma unroll
    for (int u = 0; u < kZPer; u++) {
      const int z = z0 + u * zstep;
      if (z < K3) ba[l * KS + z] = v[u];
    }
  } else {
    copy_batched<8>(
        nl * K3,
        [&](int t) {
          const int z = t / nl, l = t - z * nl;
          return spec[z * zs + l];
        },
        [&](int t, C v) {
          const int z = t / nl, l = t - z * nl;
          ba[l * KS + z] = v;
        });
  }
  tw.stage(S, 2, twz, -1, nullptr);
  __syncthreads();
  C *buf = fft_axis<C, -1>(S, 2, twz, ba, bb, nl, KS);
  if (fixed) {
    const int l = threadIdx.x % nl, z0 = threadIdx.x / nl;
#pragma unroll
    for (int u = 0; u < kZPer; u++) {
      const int z = z0 + u * zstep;
      if (z < K3) {
        C v = buf[l * KS + z];
        v.x *= g[u];
        v.y *= g[u];
        buf[l * KS + z] = v;
      }
    }
  } else {
    for (int t = threadIdx.x; t < nl * K3; t += blockDim.x) {
      const int z = t / nl, l = t - z * nl;
      const R gg = bc[z * zs + l];
      C v = buf[l * KS + z];
      v.x *= gg;
      v.y *= gg;
      buf[l * KS + z] = v;
    }
  }
  __syncthreads();
  const C *out = fft_axis<C, +1>(S, 2, twz, buf, buf == ba ? bb : ba, nl, KS);
  if (fixed) {
    const int l = threadIdx.x % nl, z0 = threadIdx.x / nl;
#pragma unroll
    for (int u = 0; u < kZPer; u++) {
      const int z = z0 + u * zstep;
      if (z < K3) spec[(size_t)z * zs + l] = out[l * KS + z];
    }
  } else {
    for (int t = threadIdx.x; t < nl * K3; t += blockDim.x) {
      const int z = t / nl, l = t - z * nl;
      spec[z * zs + l] = out[l * KS + z];
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kPlaneBlock) void k_fftz_conv(const Rep *__restrict__ reps, WMap wm) {
  extern __shared__ char smem[];
  fftz_block<R>(reps, wm, blockIdx.x, smem);
}

// One charge's reciprocal force from the potential mesh (IFrc,
// compt_acc.f90:1402-1467): F = -q sum over the p^3 points of the spline
// weight products (one of them differentiated) times the potential; rows of
// p points along x, rows summed z outer, y inner.  The order is a template
// parameter so that the weights stay in registers.
template <typename R, int O>
__device__ __forceinline__ void gather_one(const Rep &S, double4 p, int cz, R &f1, R &f2, R &f3) {
  R wx[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder], dx[kMaxOrder], dy[kMaxOrder], dz[kMaxOrder];
  const int K1 = S.K[0], K2 = S.K[1], K3 = S.K[2];
  const int bx = wrapk(spline<R, O>(p.x, S.kg[0], O, wx, dx), K1);
  const int by = wrapk(spline<R, O>(p.y, S.kg[1], O, wy, dy), K2);
  const int bz = wrapk(spline<R, O>(p.z, S.kg[2], O, wz, dz), K3);
  const R q = (R)p.w;
  int xs[O];
#pragma unroll
  for (int a = 0; a < O; a++) {
    int x = bx + a;
    xs[a] = x >= K1 ? x - K1 : x;
  }
  const R *mesh = (const R *)S.meshu;
  f1 = f2 = f3 = 0;
#pragma unroll
  for (int c = 0; c < O; c++) {
    if (c % kGatherLanes != cz) continue;  // this lane's z planes
    int z = bz + c;
    z -= z >= K3 ? K3 : 0;
#pragma unroll
    for (int b = 0; b < O; b++) {
      int y = by + b;
      y -= y >= K2 ? K2 : 0;
      const R *row = mesh + ((size_t)z * K2 + y) * K1;
      R s1 = 0, s2 = 0;
#pragma unroll
      for (int a = 0; a < O; a++) {
        const R v = row[xs[a]];
        s1 += (dx[a] * q) * v;
        s2 += (wx[a] * q) * v;
      }
      f1 += wz[c] * wy[b] * s1;
      f2 += wz[c] * dy[b] * s2;
      f3 += dz[c] * wy[b] * s2;
    }
  }
}

// The same with an order per axis (ordr(1:3) unequal), loops not unrolled.
template <typename R>
__device__ __forceinline__ void gather_axes(const Rep &S, double4 p, int cz, R &f1, R &f2, R &f3) {
  R wx[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder], dx[kMaxOrder], dy[kMaxOrder], dz[kMaxOrder];
  const int K1 = S.K[0], K2 = S.K[1], K3 = S.K[2];
  const int bx = wrapk(spline<R, 0>(p.x, S.kg[0], S.ord[0], wx, dx), K1);
  const int by = wrapk(spline<R, 0>(p.y, S.kg[1], S.ord[1], wy, dy), K2);
  const int bz = wrapk(spline<R, 0>(p.z, S.kg[2], S.ord[2], wz, dz), K3);
  const R q = (R)p.w;
  const R *mesh = (const R *)S.meshu;
  f1 = f2 = f3 = 0;
  for (int c = cz; c < S.ord[2]; c += kGatherLanes) {  // this lane's z planes
    int z = bz + c;
    z -= z >= K3 ? K3 : 0;
    for (int b = 0; b < S.ord[1]; b++) {
      int y = by + b;
      y -= y >= K2 ? K2 : 0;
      const R *row = mesh + ((size_t)z * K2 + y) * K1;
      R s1 = 0, s2 = 0;
      for (int a = 0; a < S.ord[0]; a++) {
        int x = bx + a;
        x -= x >= K1 ? K1 : 0;
        const R v = row[x];
        s1 += (dx[a] * q) * v;
        s2 += (wx[a] * q) * v;
      }
      f1 += wz[c] * wy[b] * s1;
      f2 += wz[c] * dy[b] * s2;
      f3 += dz[c] * wy[b] * s2;
    }
  }
}

// Reciprocal force of each charge (kGatherLanes lanes per charge, each
// summing the z planes c = its lane mod kGatherLanes, then a fixed DPP tree
// over the quad), plus per-block sums of the force (for ZeroForce,
// compt_acc.f90:1470-1492) and of q z lb/Beta (the slab dipole M_z,
// compt_acc.f90:1041-1044), in double.
template <typename R>
__device__ __forceinline__ void gather_block(const Rep *__restrict__ reps, WMap wm, int gb) {
  const int r = find_rep(wm, gb);
  const Rep &S = reps[r];
  const int blk = gb - wm.blk[r];
  static_assert(kGatherLanes == 4, "a charge's lanes are a quad of the DPP tree");
  const int cz = threadIdx.x & (kGatherLanes - 1);
  const int m = blk * (kBlock / kGatherLanes) + threadIdx.x / kGatherLanes;
  const bool live = m < S.Nq;
  {  // the tile (or plane) lists are consumed: clear them for the next evaluation
    const int ntile = S.nlist;
    const int nth = (wm.blk[r + 1] - wm.blk[r]) * kBlock;
    for (int k = blk * kBlock + threadIdx.x; k < ntile; k += nth) S.tcnt[k] = 0;
  }
  double f[4] = {0, 0, 0, 0};
  double4 p = make_double4(0, 0, 0, 0);
  R f1 = 0, f2 = 0, f3 = 0;
  int i = 0;
  if (live) {
    i = S.charge[m];
    p = S.pos[i];
    switch (S.order) {
      case 3: gather_one<R, 3>(S, p, cz, f1, f2, f3); break;
      case 4: gather_one<R, 4>(S, p, cz, f1, f2, f3); break;
      case 5: gather_one<R, 5>(S, p, cz, f1, f2, f3); break;
      case 6: gather_one<R, 6>(S, p, cz, f1, f2, f3); break;
      case 7: gather_one<R, 7>(S, p, cz, f1, f2, f3); break;
      case 8: gather_one<R, 8>(S, p, cz, f1, f2, f3); break;
      default: gather_axes<R>(S, p, cz, f1, f2, f3); break;
    }
  }
  f1 += dpp_move<0xb1, 0xf>(f1);  // the quad's planes: xor 1, then xor 2
  f2 += dpp_move<0xb1, 0xf>(f2);
  f3 += dpp_move<0xb1, 0xf>(f3);
  f1 += dpp_move<0x4e, 0xf>(f1);
  f2 += dpp_move<0x4e, 0xf>(f2);
  f3 += dpp_move<0x4e, 0xf>(f3);
  if (live && cz == 0) {
    f[0] = -(double)f1 * S.fnorm[0];
    f[1] = -(double)f2 * S.fnorm[1];
    f[2] = -(double)f3 * S.fnorm[2];
    f[3] = p.w * p.z * S.c1;
    S.fsp[i] = f[0];
    S.fsp[(size_t)S.NN + i] = f[1];
    S.fsp[2 * (size_t)S.NN + i] = f[2];
  }
  __shared__ double red[4][kBlock / 64];
  for (int d = 0; d < 4; d++) {
    double v = wave_sum(f[d]);
    if ((threadIdx.x & 63) == 0) red[d][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; w++) s += red[threadIdx.x][w];
    S.part[(size_t)blk * 4 + threadIdx.x] = s;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_gather(const Rep *__restrict__ reps, WMap wm) {
  gather_block<R>(reps, wm, blockIdx.x);
}

template <typename R>
__device__ __forceinline__ void store_coulomb(const Rep &S, int i, double q, const double *tot,
                                              double *out) {
  CoulParts<R> c;
  c.load(S, i);
  double ax, ay, az;
  coul_value<R>(S, c, q, tot, ax, ay, az);
  if (out) {
    out[i] = ax;
    out[(size_t)S.NN + i] = ay;
    out[2 * (size_t)S.NN + i] = az;
  } else if constexpr (sizeof(R) == 4) {
    ((float *)S.ccx)[i] = (float)ax;
    ((float *)S.ccy)[i] = (float)ay;
    ((float *)S.ccz)[i] = (float)az;
  } else {
    ((double *)S.ccx)[i] = ax;
    ((double *)S.ccy)[i] = ay;
    ((double *)S.ccz)[i] = az;
  }
}

// The real-space pair sums of charged particle i, before the lb/Beta factor.
template <typename R>
__device__ __forceinline__ void store_real(const Rep &S, int i, R fx, R fy, R fz) {
  R *f = (R *)S.freal;
  f[i] = fx;
  f[(size_t)S.NN + i] = fy;
  f[2 * (size_t)S.NN + i] = fz;
}

// Joins the two halves of a Coulomb evaluation, which run concurrently
// (SPME on one stream, the real-space sweep on another): each workgroup's
// first wavefront sums the gather kernel's block partials (mean reciprocal
// force, 3, and M_z; coul_totals, the same bits in every workgroup), and
// the workgroup stores the Coulomb force of its charges.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_coul_store(const Rep *__restrict__ reps, WMap wm,
                                                       double *__restrict__ out) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  __shared__ double tot[4];
  if (threadIdx.x < 64) {
    double t[4];
    coul_totals<4>(S, t);
    if (threadIdx.x == 0)
      for (int d = 0; d < 4; d++) tot[d] = t[d];
  }
  __syncthreads();
  const int m = (blockIdx.x - wm.blk[r]) * kBlock + threadIdx.x;
  if (m >= S.Nq) return;
  const int i = S.charge[m];
  store_coulomb<R>(S, i, S.pos[i].w, tot, out);
}

// erfc-screened Coulomb pair force of real_space (compt_acc.f90:1034-1036)
// without the lb/Beta factor: q_i q_j / r^2 (erfc(a r)/r + 2a/sqrt(pi)
// exp(-a^2 r^2)) r_ij, for r < r_c; wrap applies the x,y minimum image
// (false for an image-shifted j).
template <typename R>
__device__ __forceinline__ void real_pair(R xi, R yi, R zi, R qi, R xj, R yj, R zj, R qj, bool wrap,
                                          R Lx, R Ly, R hLx, R hLy, R rcc2, R alpha, R alpha2, R c2,
                                          R &fx, R &fy, R &fz) {
  R dx = xi - xj, dy = yi - yj, dz = zi - zj;
  if (wrap) {
    dx = image(dx, Lx, hLx);
    dy = image(dy, Ly, hLy);
  }
  const R rsq = dx * dx + dy * dy + dz * dz;
  if (rsq < rcc2) {
    const R rr = sqrt(rsq);
    const R g = qi * qj / rsq * (erfc(alpha * rr) / rr + c2 * exp(-alpha2 * rsq));
    fx += g * dx;
    fy += g * dy;
    fz += g * dz;
  }
}

// FP32 form: erfc by Abramowitz & Stegun 7.1.26, t = 1/(1 + p x),
// erfc(x) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2), absolute
// error <= 1.5e-7, which shares exp(-a^2 r^2) with the second term.  1/r,
// 1/(1 + p x) and exp2 are the hardware's single-instruction forms
// (v_rsq_f32, v_rcp_f32, v_exp_f32: about 1 ulp each); 1/r^2 = (1/r)^2.
template <>
__device__ __forceinline__ void real_pair<float>(float xi, float yi, float zi, float qi, float xj,
                                                 float yj, float zj, float qj, bool wrap, float Lx,
                                                 float Ly, float hLx, float hLy, float rcc2,
                                                 float alpha, float alpha2, float c2, float &fx,
                                                 float &fy, float &fz) {
  float dx = xi - xj, dy = yi - yj, dz = zi - zj;
  if (wrap) {
    dx = image(dx, Lx, hLx);
    dy = image(dy, Ly, hLy);
  }
  const float rsq = dx * dx + dy * dy + dz * dz;
  if (rsq < rcc2) {
    const float ir = __builtin_amdgcn_rsqf(rsq);
    const float rr = rsq * ir;
    const float e = __builtin_amdgcn_exp2f(-(alpha2 * 1.44269504f) * rsq);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * alpha * rr);
    const float poly =
        t * (0.254829592f +
             t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    const float g = qi * qj * (ir * ir) * e * (poly * ir + c2);
    fx += g * dx;
    fy += g * dy;
    fz += g * dz;
  }
}

// real_space over the reference-form list (compt_acc.f90:1012-1038) for
// replicas whose list cells number fewer than three along x or y
// (Rep::real_list), where the reference lists a neighbour cell's pairs more
// than once and sums each repeat: a wavefront per charge slot (the replica's
// wavefronts stride over its charges), its lanes taking every 64th entry of
// the list and keeping the pairs within r_c, summed by a fixed DPP tree.
// The lb/Beta factor and the slab term are applied by the Coulomb store.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_real_list(const Rep *__restrict__ reps, WMap wm) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int blk = blockIdx.x - wm.blk[r];
  const int nwave = (wm.blk[r + 1] - wm.blk[r]) * (kBlock / 64);
  const int lane = threadIdx.x & 63;
  const R Lx = (R)S.Lx, Ly = (R)S.Ly, hLx = (R)S.hLx, hLy = (R)S.hLy, rcc2 = (R)S.rcc2,
          alpha = (R)S.alpha, alpha2 = (R)S.alpha2, c2 = (R)S.c2;
  for (int t = blk * (kBlock / 64) + (threadIdx.x >> 6); t < S.Nq; t += nwave) {
    const int i = S.charge[t];
    const int n = min(S.nbn2[t], S.nbcap2);
    R xi, yi, zi, qi;
    if constexpr (sizeof(R) == 4) {
      const float4 p = S.pos32[i];
      xi = p.x; yi = p.y; zi = p.z; qi = p.w;
    } else {
      const double4 p = S.pos[i];
      xi = p.x; yi = p.y; zi = p.z; qi = p.w;
    }
    R fx = 0, fy = 0, fz = 0;
    for (int k = lane; k < n; k += 64) {
      const int j = S.nb2[(size_t)k * S.Nq + t];
      if constexpr (sizeof(R) == 4) {
        const float4 p = S.pos32[j];
        real_pair<R>(xi, yi, zi, qi, p.x, p.y, p.z, p.w, true, Lx, Ly, hLx, hLy, rcc2, alpha, alpha2,
                     c2, fx, fy, fz);
      } else {
        const double4 p = S.pos[j];
        real_pair<R>(xi, yi, zi, qi, p.x, p.y, p.z, p.w, true, Lx, Ly, hLx, hLy, rcc2, alpha, alpha2,
                     c2, fx, fy, fz);
      }
    }
    fx = wave_total(fx);
    fy = wave_total(fy);
    fz = wave_total(fz);
    if (lane == 0) store_real<R>(S, i, fx, fy, fz);
  }
}

// Charges into the lists of the spreading tiles that hold their spline
// points (one to eight) and into the cells of the real-space sweep.
__global__ __launch_bounds__(kBlock) void k_bin_q(const Rep *__restrict__ reps, WMap wm, int planes) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  desc_prefetch(S);
  const int m = (blockIdx.x - wm.blk[r]) * kBlock + threadIdx.x;
  const bool live = m < S.Nq;
  int i = 0;
  double4 p = make_double4(0, 0, 0, 0);
  if (live) {
    i = S.charge[m];
    p = S.pos[i];
  }
  int t0[3] = {0, 0, 0}, nt[3] = {0, 0, 0}, bz = 0;
  if (live && planes) {  // plane spreading (k_plane_spread): one list per first z index
    bz = spline_base(p.z, S.kg[2], S.ord[2], S.K[2]);
  } else if (live) {  // the spreading tiles of the charge's spline points
    const double c3[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int o = S.ord[d];
      const int b = spline_base(c3[d], S.kg[d], o, S.K[d]);
      t0[d] = tile_of(S, d, b);
      int e = b + o - 1;
      e -= e >= S.K[d] ? S.K[d] : 0;
      nt[d] = tile_of(S, d, e) != t0[d] ? 2 : 1;
    }
  }
  // nine appends (up to eight tiles, the sweep cell): the runs first, then
  // every atomic issued before any result is used, so that their latencies
  // overlap
  int key[9];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int a = k & 1, b = (k >> 1) & 1, c = k >> 2;
    key[k] = -1;
    if (a < nt[0] && b < nt[1] && c < nt[2]) {
      const int tx = (t0[0] + a) % S.ntl[0], ty = (t0[1] + b) % S.ntl[1],
                tz = (t0[2] + c) % S.ntl[2];
      // sub-list by groups of eight consecutive charges
      key[k] = ((tz * S.ntl[1] + ty) * S.ntl[0] + tx) * kTileSub + ((m >> 3) & (kTileSub - 1));
    }
  }
  if (live && planes) key[0] = bz * kTileSub + ((m >> 3) & (kTileSub - 1));
  key[8] = -1;
  if (live && S.real_cells) {
    int a, b, c;
    if (cell_of(p, S.nc3, S.rcel3, S.hLx, S.hLy, a, b, c)) key[8] = (c * S.nc3[1] + b) * S.nc3[0] + a;
    else set_error(S.ctrl, MD_ECELL, i, -1);
  }
  // the appends in use: one plane list (plane spreading) or up to eight
  // tiles, and the sweep cell; the keys of the others are empty in every
  // lane (uniform: planes and real_cells are the launch's and the
  // replica's), so their run detection and slot broadcast are skipped
  const int nk = planes ? 1 : 8;
  const bool cells = S.real_cells != 0;
  int *ctr[9];
  WaveRun run[9];
  int base[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    ctr[k] = key[k] < 0 ? nullptr : (k < 8 ? S.tcnt : S.cnt3) + key[k];
    run[k] = WaveRun{0, 0, false};
    if (k < nk || (k == 8 && cells)) run[k] = wave_run(ctr[k]);
  }
#pragma unroll
  for (int k = 0; k < 9; k++) base[k] = run[k].head && ctr[k] ? atomicAdd(ctr[k], run[k].n) : 0;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    if (!(k < nk || (k == 8 && cells))) continue;
    const int slot = __shfl(base[k], run[k].h, 64) + (lane - run[k].h);
    if (!ctr[k]) continue;
    if (k < 8) {
      if (slot < S.capt) S.tmem[(size_t)key[k] * S.capt + slot] = i;
      else atomicOr(&S.ctrl->ovf, 32);
    } else {
      if (slot < S.ccap3) S.mem3[(size_t)key[k] * S.ccap3 + slot] = i;
      else atomicOr(&S.ctrl->ovf, 16);
    }
  }
}

// One workgroup per cell: sort the members, keep the sorted count, clear
// the binning count.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_sort_q(const Rep *__restrict__ reps, WMap wm) {
  using V = typename std::conditional<sizeof(R) == 4, float4, double4>::type;
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  const int cell = blockIdx.x - wm.blk[r];
  const int raw = S.cnt3[cell];
  const int n = min(raw, S.ccap3);
  extern __shared__ int sbuf[];
  int *mem = S.mem3 + (size_t)cell * S.ccap3;
  // each member's place in ascending order is the number of members below
  // it (the members are distinct particles, whose order is the charge
  // slots' order: md_build_topology lists the charges by ascending particle
  // index): one pass over the cell's members
  // in LDS (broadcast reads), then the member and its position go straight
  // to their sorted slots
  for (int k = threadIdx.x; k < n; k += blockDim.x) sbuf[k] = mem[k];
  __syncthreads();
  V *ps = (V *)S.pos3s + (size_t)cell * S.ccap3;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int i = sbuf[k];
    int rank = 0;
    for (int j = 0; j < n; j++) rank += sbuf[j] < i;
    mem[rank] = i;
    if constexpr (sizeof(R) == 4) ps[rank] = S.pos32[i];
    else ps[rank] = S.pos[i];
  }
  if (threadIdx.x == 0) {
    S.scnt3[cell] = n;
    S.cnt3[cell] = 0;
    // the largest occupancy seen (read first: one address for every cell)
    if (raw > S.ctrl->max_cell2) atomicMax(&S.ctrl->max_cell2, raw);
  }
}


// Candidates staged in LDS per pass of a k_real_cells wavefront.
template <typename R> struct RealTile { static constexpr int N = 256; };
template <> struct RealTile<float> { static constexpr int N = 512; };
static_assert(RealTile<float>::N % kRealRound == 0 && RealTile<double>::N % kRealRound == 0,
              "tiles are padded to whole compaction rounds");

// LDS layout of the sweep's workgroup form (real_cells_wg) in workgroups of
// BS lanes: the candidate tile, a hit list per wavefront, the item's target
// sums (double), the stencil offsets and cells, the work list, the
// wavefronts' counts.
template <typename R, int BS>
struct RealWG {
  using V = typename std::conditional<sizeof(R) == 4, float4, double4>::type;
  static constexpr int TN = RealTile<R>::N, NW = BS / 64;
  static constexpr size_t cand = 0;
  static constexpr size_t hits = cand + (size_t)TN * sizeof(V);
  static constexpr size_t accs = hits + (size_t)NW * TN * sizeof(short);
  static constexpr size_t offs = accs + (size_t)kRealChunkWG * 3 * sizeof(double);
  static constexpr size_t cids = offs + 28 * sizeof(int);
  static constexpr size_t work = cids + 28 * sizeof(int);
  static constexpr size_t wcnt = work + (size_t)BS * sizeof(int);
  static constexpr size_t bytes = (wcnt + (size_t)NW * sizeof(int) + 15) / 16 * 16;
  static_assert(accs % 8 == 0, "the target sums are doubles");
};

// Periodic image of stencil slot l's members for a target in cell (a, b),
// with three or more cells of edge >= r_c along x and y: the x, y offset
// that places them beside the target.  Each stencil cell is then a distinct
// direct neighbour, and the shifted copy gives every pair within r_c its
// minimum-image separation (|dx| < r_c <= L/3; any other image is at least
// L - 2e >= e >= r_c away for cell edge e <= L/3), while a pair the minimum
// image puts at r_c or beyond stays there.  The sweep stages shifted
// candidates and tests them without the minimum image.
template <typename R>
__device__ __forceinline__ void slot_shift(int l, int a, int b, const int nc[3], R Lx, R Ly, R &sx,
                                           R &sy) {
  const int p = a + l / 9 - 1, q = b + (l / 3) % 3 - 1;
  sx = p < 0 ? -Lx : (p >= nc[0] ? Lx : (R)0);
  sy = q < 0 ? -Ly : (q >= nc[1] ? Ly : (R)0);
}

// Whether the sweep stages image-shifted candidates (slot_shift) for a
// replica: three or more real-space cells along x and y.
__device__ __forceinline__ bool shifted_stencil(const int nc[3]) { return nc[0] >= 3 && nc[1] >= 3; }

// Sentinel of the real-space tiles past their last candidate: far beyond
// r_c of any target (its squared distance stays finite), charge 0.
template <typename V> __device__ __forceinline__ V tile_pad() {
  using R = decltype(V{}.x);
  constexpr R far = sizeof(R) == 4 ? (R)1e18 : (R)1e150;
  return V{far, far, far, (R)0};
}

// Force on one target from a staged tile of nc candidates (candidate `self`
// is the target itself), by one wavefront: the candidates inside r_c are
// compacted in candidate order into hit[], the lanes stride over them, and
// a fixed DPP tree sums the lanes (wave_total).  The tile is padded with
// far-away sentinels (tile_pad) up to the next multiple of kRealRound
// candidates.  wrap = false for a tile
// staged image-shifted (slot_shift): the tests and pair forces then take the
// separations as they stand; wrap = true (fewer than three cells along x or
// y) applies the minimum image.  Both forms of the sweep use it with the
// same tiles, so a target's sum does not depend on the form.
template <typename R, typename V>
__device__ __forceinline__ void target_tile(const V pi, const V *cand, bool wrap, int nc, int self,
                                            short *hit, R Lx, R Ly, R hLx, R hLy, R rcc2, R alpha,
                                            R alpha2, R c2, R &fx, R &fy, R &fz) {
  const int lane = threadIdx.x & 63;
  // the compaction keeps a superset of the pairs within r_c (its radius is
  // a few roundings wider); real_pair tests each exactly
  const R rin2 = rcc2 * (R)(sizeof(R) == 4 ? 1.00001 : 1.000000000001);
  int nh = 0;
  // four rounds of 64 candidates at a time: their LDS loads are issued
  // together; the hits are appended round by round, in candidate order
  constexpr int U = kRealRound / 64;
  for (int k0 = 0; k0 < nc; k0 += kRealRound) {
    V pj[U];
#pragma unroll
    for (int u = 0; u < U; u++) pj[u] = cand[k0 + u * 64 + lane];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = k0 + u * 64 + lane;
      R dx = pi.x - pj[u].x, dy = pi.y - pj[u].y;
      const R dz = pi.z - pj[u].z;
      if (wrap) {
        dx = image(dx, Lx, hLx);
        dy = image(dy, Ly, hLy);
      }
      in[u] = k != self && dx * dx + dy * dy + dz * dz < rin2;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const unsigned long long mk = __builtin_amdgcn_ballot_w64(in[u]);
      // hits of lower lanes: v_mbcnt over the ballot
      const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(mk >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((unsigned)mk, 0u));
      if (in[u]) hit[nh + below] = (short)(k0 + u * 64 + lane);
      nh += __popcll(mk);
    }
  }
  wave_sync();
  fx = fy = fz = 0;
  for (int h = lane; h < nh; h += 64) {
    const V pj = cand[hit[h]];
    real_pair<R>(pi.x, pi.y, pi.z, pi.w, pj.x, pj.y, pj.z, pj.w, wrap, Lx, Ly, hLx, hLy, rcc2, alpha,
                 alpha2, c2, fx, fy, fz);
  }
  wave_sync();  // hit[] is refilled by the next target
  fx = wave_total(fx);
  fy = wave_total(fy);
  fz = wave_total(fz);
}

// The sweep form for engines with few charges: a work item is a chunk of
// kRealChunkWG targets of one cell, taken by a whole workgroup.  The 27
// stencil cells' members are staged once per tile of RealTile<R>::N
// candidates by all kBlock lanes; the wavefronts share the item's targets
// (target_tile each) and sum them over the tiles in LDS (double), in the
// same order as the wavefront form.  Workgroups stride over the items
// (chunk-major, as in the wavefront form) and skip the empty ones in bulk.
template <typename R, typename V, int BS>
__device__ void real_cells_wg(const Rep &S, int b0, int nb, char *lds, int part, int nparts) {
  constexpr int TN = RealTile<R>::N, NW = BS / 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nch = (S.ccap3 + kRealChunkWG - 1) / kRealChunkWG;
  const int ncl = S.nc3[0] * S.nc3[1] * S.nc3[2], nitem = ncl * nch;
  // the workgroup's LDS (RealWG<R, BS>): the candidate tile, a hit list per
  // wavefront, the target sums, the stencil table and the work list; no
  // static LDS, so that kernels sharing a launch keep their dynamic LDS
  // where they expect it
  using L = RealWG<R, BS>;
  V *cand = (V *)(lds + L::cand);
  short *hit = (short *)(lds + L::hits) + w * TN;
  double (*accs)[3] = (double(*)[3])(lds + L::accs);
  int *offs = (int *)(lds + L::offs), *cids = (int *)(lds + L::cids);
  int *work = (int *)(lds + L::work), *wcnt = (int *)(lds + L::wcnt);
  const V *ps = (const V *)S.pos3s;
  const R Lx = (R)S.Lx, Ly = (R)S.Ly, hLx = (R)S.hLx, hLy = (R)S.hLy, rcc2 = (R)S.rcc2;
  const bool wrap = !shifted_stencil(S.nc3);  // else the tiles are staged image-shifted
  const R alpha = (R)S.alpha, alpha2 = (R)S.alpha2, c2 = (R)S.c2;
  // the workgroup's items are b0 + j nb; this launch takes those with
  // j = part mod nparts, kBlock at a time, the non-empty ones compacted in order
  const int step = nb * nparts;
  for (int it0 = b0 + part * nb; it0 < nitem; it0 += step * BS) {
    const int item = it0 + threadIdx.x * step;
    bool ne = false;
    if (item < nitem) {
      const int cell = item % ncl;
      ne = S.scnt3[cell] > (item / ncl) * kRealChunkWG;
    }
    const unsigned long long bal = __ballot(ne);
    if (lane == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    int off = 0, nwork = 0;
    for (int k = 0; k < NW; k++) {
      off += k < w ? wcnt[k] : 0;
      nwork += wcnt[k];
    }
    if (ne) work[off + __popcll(bal & ((1ull << lane) - 1ull))] = item;
    __syncthreads();
    for (int q = 0; q < nwork; q++) {
      const int it = work[q];
      const int cell = it % ncl, t0 = (it / ncl) * kRealChunkWG;
      const int nt = min(S.scnt3[cell] - t0, kRealChunkWG);
      const int ca = cell % S.nc3[0], cb = (cell / S.nc3[0]) % S.nc3[1];
      const int total = stencil_offsets(cell, S.nc3, S.scnt3, offs, cids);
      // stencil slot 13 is the cell itself: candidate offs[13] + t0 + tg is target tg
      const int self0 = offs[13] + t0;
      const V *tp = ps + (size_t)cell * S.ccap3 + t0;
      // the item's targets, lane tg of every wavefront holding target tg:
      // one load per item, read out per target (readlane), instead of a
      // dependent global load before each target's tile
      static_assert(kRealChunkWG <= 64, "an item's targets fit one wavefront's lanes");
      V pt{};
      if (lane < nt) pt = tp[lane];
      if (threadIdx.x < 3 * kRealChunkWG) accs[threadIdx.x / 3][threadIdx.x % 3] = 0.0;
      for (int base = 0; base < total; base += TN) {
        const int nc = min(TN, total - base);
        const int ncp = (nc + kRealRound - 1) / kRealRound * kRealRound;
        for (int k = threadIdx.x; k < ncp; k += BS) {
          V c = tile_pad<V>();
          if (k < nc) {
            const int t = base + k, l = stencil_slot(offs, t);
            c = ps[(size_t)cids[l] * S.ccap3 + (t - offs[l])];
            if (!wrap) {
              R sx, sy;
              slot_shift(l, ca, cb, S.nc3, Lx, Ly, sx, sy);
              c.x += sx;
              c.y += sy;
            }
          }
          cand[k] = c;
        }
        __syncthreads();
        for (int tg = w; tg < nt; tg += NW) {
          V pi;
          pi.x = lane_value(pt.x, tg);
          pi.y = lane_value(pt.y, tg);
          pi.z = lane_value(pt.z, tg);
          pi.w = lane_value(pt.w, tg);
          R fx, fy, fz;
          target_tile<R>(pi, cand, wrap, nc, self0 + tg - base, hit, Lx, Ly, hLx, hLy, rcc2, alpha,
                         alpha2, c2, fx, fy, fz);
          if (lane == 0) {
            accs[tg][0] += (double)fx;
            accs[tg][1] += (double)fy;
            accs[tg][2] += (double)fz;
          }
        }
        __syncthreads();
      }
      if (threadIdx.x < nt)
        store_real<R>(S, S.mem3[(size_t)cell * S.ccap3 + t0 + threadIdx.x], (R)accs[threadIdx.x][0],
                      (R)accs[threadIdx.x][1], (R)accs[threadIdx.x][2]);
      __syncthreads();  // accs, offs and the tile are reused
    }
    __syncthreads();  // work[] is refilled
  }
}

// real_space (compt_acc.f90:980-1049) by a cell sweep over cells of edge
// >= r_c: a work item is a chunk of kRealChunk target charges of one cell,
// and each wavefront works through its own items (engines with few charges,
// whose items are too few to fill the chip, take real_cells_wg instead).  The members of the
// cell's 27 neighbour cells (reference stencil order, ascending within a
// cell) are staged in the wavefront's LDS tile from the cell-sorted position
// copy, image-shifted (slot_shift); for each target the candidates inside
// r_c are compacted first, the 64 lanes stride over those and a DPP tree
// (fixed order) sums the lanes; lane tg keeps target tg's
// sum over the tiles and the lanes store their targets together.
// Wavefronts stride over the items and skip the empty ones 64 at a time.
template <typename R>
struct RealLDS {  // LDS of a real-space workgroup of kBlock lanes, either form:
                  // the wavefront form's candidate tiles and hit lists
  using V = typename std::conditional<sizeof(R) == 4, float4, double4>::type;
  static constexpr int NW = kBlock / 64, TN = RealTile<R>::N;
  static constexpr size_t wave = (size_t)NW * TN * (sizeof(V) + sizeof(short));
  static constexpr size_t bytes = wave > RealWG<R, kBlock>::bytes ? wave : RealWG<R, kBlock>::bytes;
};

// Real-space workgroup gb of the sweep's block map (its LDS at lds: at
// least RealLDS<R>::bytes, or RealWG<R, BS>::bytes for the workgroup form).
template <typename R, bool WG, int BS = kBlock>
__device__ __forceinline__ void real_cells_block(const Rep *__restrict__ reps, WMap wm, int gb,
                                                 char *lds, int part = 0, int nparts = 1) {
  using V = typename RealLDS<R>::V;
  static_assert(WG || BS == kBlock, "the wavefront form runs in kBlock workgroups");
  constexpr int TN = RealTile<R>::N, NW = BS / 64;
  const int r = find_rep(wm, gb);
  const Rep &S = reps[r];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = (gb - wm.blk[r]) * NW + w, nwv = (wm.blk[r + 1] - wm.blk[r]) * NW;
  // item = chunk * ncell + cell (every cell's first chunk, then every
  // cell's second, ...: the non-empty items are dense at the front, so that
  // the strided assignment deals them out evenly); wavefront gw takes items
  // gw, gw + nwv, ...
  const int nch = (S.ccap3 + kRealChunk - 1) / kRealChunk;
  const int ncl = S.nc3[0] * S.nc3[1] * S.nc3[2], nitem = ncl * nch;
  V *cand_all = (V *)lds;
  short *hit_all = (short *)(lds + (size_t)NW * TN * sizeof(V));
  V *cand = cand_all + w * TN;
  short *hit = hit_all + w * TN;
  const V *ps = (const V *)S.pos3s;
  const R Lx = (R)S.Lx, Ly = (R)S.Ly, hLx = (R)S.hLx, hLy = (R)S.hLy, rcc2 = (R)S.rcc2;
  const bool wrap = !shifted_stencil(S.nc3);  // else the tiles are staged image-shifted
  const R alpha = (R)S.alpha, alpha2 = (R)S.alpha2, c2 = (R)S.c2;
  if constexpr (WG) {
    real_cells_wg<R, V, BS>(S, gb - wm.blk[r], wm.blk[r + 1] - wm.blk[r], lds, part, nparts);
    return;
  }
  // the wavefront's items are gw + j nwv; this launch takes j = part mod nparts
  const int step = nwv * nparts;
  for (int it0 = gw + part * nwv; it0 < nitem; it0 += step * 64) {
    // this wavefront's next 64 items
    const int item = it0 + lane * step;
    bool ne = false;
    if (item < nitem) {
      const int cell = item % ncl;
      ne = S.scnt3[cell] > (item / ncl) * kRealChunk;
    }
    unsigned long long pend = __ballot(ne);
    while (pend) {
      const int src = __ffsll((long long)pend) - 1;
      pend &= pend - 1;
      const int it = __shfl(item, src, 64);
      const int cell = it % ncl, t0 = (it / ncl) * kRealChunk;
      const int nt = min(S.scnt3[cell] - t0, kRealChunk);
      const int ca = cell % S.nc3[0], cb = (cell / S.nc3[0]) % S.nc3[1],
                cc = cell / (S.nc3[0] * S.nc3[1]);
      // stencil cell of lane l < 27, its count and exclusive offset
      int cl = -1, cn = 0;
      if (lane < 27) {
        cl = stencil_cell<true>(lane, ca, cb, cc, S.nc3);
        cn = cl >= 0 ? S.scnt3[cl] : 0;
      }
      int incl = cn;
      for (int o = 1; o < 32; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      const int excl = incl - cn;
      const int total = __shfl(incl, 26, 64);
      // stencil slot 13 is the cell itself: candidate self0 + tg is target tg
      const int self0 = __shfl(excl, 13, 64) + t0;
      // lane tg holds target tg's position and accumulates its force
      V pt{};
      if (lane < nt) pt = ps[(size_t)cell * S.ccap3 + t0 + lane];
      double ax = 0, ay = 0, az = 0;
      for (int base = 0; base < total; base += TN) {
        const int nc = min(TN, total - base);
        // stage the tile: TN/64 candidates per lane, loads issued together
        constexpr int U = TN / 64;
        size_t src[U];
        int sl[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int t = base + u * 64 + lane;
          int l = 0;  // the largest stencil slot with excl_l <= t
#pragma unroll
          for (int st = 16; st > 0; st >>= 1) {
            const int c = l + st;
            const int ex = __shfl(excl, c < 27 ? c : 26, 64);
            if (c < 27 && ex <= t) l = c;
          }
          const int lc = __shfl(cl, l, 64), lx = __shfl(excl, l, 64);
          src[u] = (size_t)lc * S.ccap3 + (t - lx);
          sl[u] = l;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {  // the whole tile: candidates, then sentinels
          V c = tile_pad<V>();
          if (u * 64 + lane < nc) {
            c = ps[src[u]];
            if (!wrap) {
              R sx, sy;
              slot_shift(sl[u], ca, cb, S.nc3, Lx, Ly, sx, sy);
              c.x += sx;
              c.y += sy;
            }
          }
          cand[u * 64 + lane] = c;
        }
        wave_sync();
        for (int tg = 0; tg < nt; tg++) {
          V pi;
          pi.x = lane_value(pt.x, tg);
          pi.y = lane_value(pt.y, tg);
          pi.z = lane_value(pt.z, tg);
          pi.w = lane_value(pt.w, tg);
          R fx, fy, fz;
          target_tile<R>(pi, cand, wrap, nc, self0 + tg - base, hit, Lx, Ly, hLx, hLy, rcc2, alpha,
                         alpha2, c2, fx, fy, fz);
          if (lane == tg) {
            ax += (double)fx;
            ay += (double)fy;
            az += (double)fz;
          }
        }
        wave_sync();  // the tile is refilled
      }
      if (lane < nt) store_real<R>(S, S.mem3[(size_t)cell * S.ccap3 + t0 + lane], (R)ax, (R)ay, (R)az);
    }
  }
}

template <typename R, bool WG>
__global__ __launch_bounds__(kBlock) void k_real_cells(const Rep *__restrict__ reps, WMap wm) {
  __shared__ __align__(16) char lds[RealLDS<R>::bytes];
  real_cells_block<R, WG>(reps, wm, blockIdx.x, lds);
}

// The real-space sweep shares launches with the SPME kernels that leave the
// chip idle: workgroups past the spreading tiles (or past the gather's
// charge blocks) are the sweep's workgroups, each doing every other of its
// items (part 0 with the spreading, part 1 with the gather; a workgroup's
// duration is set by its own items, so each launch takes about half), in one
// LDS buffer sized for either kind.  Every item is computed as in the
// sweep's own launch, so the sums are the same.
template <typename R, bool WG>
__global__ __launch_bounds__(kBlock) void k_spread_real(const Rep *__restrict__ reps, WMap ws,
                                                        WMap wr) {
  constexpr size_t B = RealLDS<R>::bytes > kTile * kTile * kTile * 8 ? RealLDS<R>::bytes
                                                                     : kTile * kTile * kTile * 8;
  __shared__ __align__(16) char lds[B];
  if ((int)blockIdx.x < ws.total) spread_block<R>(reps, ws, blockIdx.x, (unsigned long long *)lds);
  else real_cells_block<R, WG>(reps, wr, (int)blockIdx.x - ws.total, lds, 0, 2);
}

template <typename R, bool WG>
__global__ __launch_bounds__(kBlock) void k_gather_real(const Rep *__restrict__ reps, WMap wg,
                                                        WMap wr, int nparts, int lead) {
  __shared__ __align__(16) char lds[RealLDS<R>::bytes];
  const int b = blockIdx.x;  // riders 0..lead-1, the gather's workgroups, the other riders
  if (b >= lead && b < lead + wg.total) gather_block<R>(reps, wg, b - lead);
  else real_cells_block<R, WG>(reps, wr, b < lead ? b : b - wg.total, lds, nparts - 1, nparts);
}

// With plane spreading the sweep's workgroup form (in workgroups of
// kPlaneBlock lanes, its LDS in the launch's dynamic LDS) can also ride in
// the plane-spreading launch (part 0) and the z-pass launch (part 1 of 3);
// the gather launch takes the last part.  Each of these workgroups runs
// one kind only.
template <typename R>
__global__ __launch_bounds__(kPlaneBlock) void k_plane_spread_real(const Rep *__restrict__ reps,
                                                                   WMap wp, WMap wr, int nparts,
                                                                   int store_q, int lead) {
  extern __shared__ char smem[];
  const int b = blockIdx.x;  // riders 0..lead-1, the plane workgroups, the other riders
  if (b >= lead && b < lead + wp.total) plane_spread_block<R>(reps, wp, b - lead, smem, store_q);
  else real_cells_block<R, true, kPlaneBlock>(reps, wr, b < lead ? b : b - wp.total, smem, 0, nparts);
}

template <typename R>
__global__ __launch_bounds__(kPlaneBlock) void k_fftz_real(const Rep *__restrict__ reps, WMap wz,
                                                           WMap wr, int nparts, int lead) {
  extern __shared__ char smem[];
  const int b = blockIdx.x;  // riders 0..lead-1, the z-pass workgroups, the other riders
  if (b >= lead && b < lead + wz.total) fftz_block<R>(reps, wz, b - lead, smem);
  else real_cells_block<R, true, kPlaneBlock>(reps, wr, b < lead ? b : b - wz.total, smem, 1, nparts);
}

// ---------------------------------------------------- rescale_velocity

// Whether a replica of a batch rescales after the step just completed:
// the pre-heat rule of main.f90:67-75 (every 20 steps below 300000, up to
// its last pre-heat step).  The list kernel has already advanced the step.
__device__ __forceinline__ bool preheat_due(const Rep &S) {
  const int last = S.ctrl->step - 1;
  return last % 20 == 0 && last <= S.preheat_end && last < 300000;
}

// rescale_velocity (init_update.f90:498-552) in three passes with fixed-order
// block sums: mean velocity over NN-N_anchor, mean square over the moving
// particles, rescale to 3/Beta.  gated: batch pre-heat, per replica.

__global__ __launch_bounds__(kBlock) void k_vsum(const Rep *__restrict__ reps, WMap wm, double *part,
                                                 int pass, int gated) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  if (gated && !preheat_due(S)) return;
  const int blk = blockIdx.x - wm.blk[r];
  const int i = blk * kBlock + threadIdx.x;
  double *pr = part + (size_t)wm.blk[r] * 4;
  const int nb = wm.blk[r + 1] - wm.blk[r];
  __shared__ double mean[3];
  if (pass == 1) {
    if (threadIdx.x < 3) {
      double s = 0;
      for (int b = 0; b < nb; b++) s += pr[(size_t)b * 4 + threadIdx.x];
      mean[threadIdx.x] = s / (S.NN - S.N_anchor);
    }
    __syncthreads();
  }
  double v[4] = {0, 0, 0, 0};
  if (i < S.NN) {
    if (pass == 0) {
      v[0] = S.vx[i];
      v[1] = S.vy[i];
      v[2] = S.vz[i];
    } else if (!(S.flags[i] & 1)) {
      const double x = S.vx[i] - mean[0], y = S.vy[i] - mean[1], z = S.vz[i] - mean[2];
      v[3] = x * x + y * y + z * z;
    }
  }
  __shared__ double red[4][kBlock / 64];
  __syncthreads();
  for (int d = 0; d < 4; d++) {
    double s = wave_sum(v[d]);
    if ((threadIdx.x & 63) == 0) red[d][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  // pass 0 writes slots 0..2 of its block; pass 1 writes slot 3 (keeping the
  // sums of pass 0 intact for the apply kernel).
  if (pass == 0 && threadIdx.x < 3) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; w++) s += red[threadIdx.x][w];
    pr[(size_t)blk * 4 + threadIdx.x] = s;
  }
  if (pass == 1 && threadIdx.x == 3) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; w++) s += red[3][w];
    pr[(size_t)blk * 4 + 3] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_vscale(const Rep *__restrict__ reps, WMap wm,
                                                   const double *part, int gated, int redrift) {
  const int r = find_rep(wm, blockIdx.x);
  const Rep &S = reps[r];
  if (gated && !preheat_due(S)) return;
  const int i = (blockIdx.x - wm.blk[r]) * kBlock + threadIdx.x;
  const double *pr = part + (size_t)wm.blk[r] * 4;
  const int nb = wm.blk[r + 1] - wm.blk[r];
  __shared__ double sh[4];
  if (threadIdx.x < 4) {
    double s = 0;
    for (int b = 0; b < nb; b++) s += pr[(size_t)b * 4 + threadIdx.x];
    sh[threadIdx.x] = s;
  }
  __syncthreads();
  double d1 = 0, d2 = 0;
  if (i < S.NN) {
    double vx = 0, vy = 0, vz = 0;  // anchors stay at rest
    if (!(S.flags[i] & 1)) {
      const double n = S.NN - S.N_anchor;
      const double v2 = sh[3] / (S.NN - S.N_anchor - 1);
      const double s = sqrt(3.0 / S.Beta / v2);
      vx = (S.vx[i] - sh[0] / n) * s;
      vy = (S.vy[i] - sh[1] / n) * s;
      vz = (S.vz[i] - sh[2] / n) * s;
    }
    S.vx[i] = vx;
    S.vy[i] = vy;
    S.vz[i] = vz;
    if (redrift) {
      double4 p = S.pos[i];
      const double dr = drift(S, p, half_kick(S, S.ax[i], vx), half_kick(S, S.ay[i], vy),
                              half_kick(S, S.az[i], vz));
      S.pos_nx[i] = p;
      if (S.pos32_nx) S.pos32_nx[i] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
      d1 = dr;
      if (p.w != 0) d2 = dr;
    }
  }
  if (redrift)
    dmax_group(S, S.par ^ 1, (blockIdx.x - wm.blk[r]) * (kBlock / 64) + (threadIdx.x >> 6), d1, d2);
}

// ------------------------------------------------------- observables

// Brush height and z-density of the polymer beads: per-block sums of z
// (height = mean z over the Npe beads, in_output.f90:1017-1020) and a
// histogram of SizeHist bins of Lz/SizeHist with ceiling binning
// (in_output.f90:1878-1886).  Integer bins and fixed-order block sums keep
// the result independent of scheduling.
__global__ void k_hist_clear(int *hist) {
  for (int k = threadIdx.x; k < kHistBins; k += blockDim.x) hist[k] = 0;
}

__global__ __launch_bounds__(kBlock) void k_observe(const Rep *__restrict__ reps, int *hist,
                                                    double *hsum) {
  const Rep &S = reps[0];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double z = 0;
  if (i < S.Npe) {
    z = S.pos[i].z;
    const int k = (int)ceil(z / (S.Lz / kHistBins));
    if (k >= 1 && k <= kHistBins) atomicAdd(&hist[k - 1], 1);
  }
  __shared__ double red[kBlock / 64];
  z = wave_sum(z);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = z;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; w++) s += red[w];
    hsum[blockIdx.x] = s;
  }
}

// Per-molecule heights of the height analysis (in_output.f90:1017-1079),
// one thread per molecule, written to mol[4 k ..]: linear chain c (its Nml
// beads from the grafted one): sum, maximum and end z; star s (the
// arm*Nma+1 beads of one molecule: the grafted arm up to the branch point,
// then arms 2..arm): sum and maximum z, branch-point z, and the sum of the
// ends of arms 2..arm.  The host adds them up in molecule order.
__global__ __launch_bounds__(kBlock) void k_heights(const Rep *__restrict__ reps, double *mol) {
  const Rep &S = reps[0];
  const int per = S.arm * S.Nma + 1, nta = S.Nga * per;
  for (int k = threadIdx.x; k < S.Nga + S.Ngl; k += blockDim.x) {
    double sum = 0, mx = 0, a = 0, b = 0;
    if (k < S.Ngl) {
      const int i0 = nta + k * S.Nml;
      for (int j = 0; j < S.Nml; j++) {
        const double z = S.pos[i0 + j].z;
        sum += z;
        if (mx < z) mx = z;
      }
      a = S.pos[i0 + S.Nml - 1].z;
    } else {
      const int i0 = (k - S.Ngl) * per;
      for (int j = 0; j < per; j++) {
        const double z = S.pos[i0 + j].z;
        sum += z;
        if (mx < z) mx = z;
      }
      a = S.pos[i0 + S.Nma].z;  // the branch point
      for (int j = 2; j <= S.arm; j++) b += S.pos[i0 + j * S.Nma].z;
    }
    double *o = mol + 4 * (size_t)k;
    o[0] = sum;
    o[1] = mx;
    o[2] = a;
    o[3] = b;
  }
}

// ------------------------------------------------------------- energies

// Energies and pair virials of one replica (reps[0]) at its current state,
// in FP64 sums, with the pair and bond arithmetic in the engine's precision
// R (as its forces).  The reference computes only the kinetic energy
// (in_output.f90:1307-1311); the potentials are those whose negative
// gradients are its forces: WCA 4((s/r)^12 - (s/r)^6) + 1 within the class
// cut (compt_acc.f90:922-952), the wall 0.4/3 ((s/d)^9 - (s/d)^3) where its
// force acts (:955-975), FENE -k R0^2/2 ln(1 - r^2/R0^2) (:865-892), the
// screened real-space pair q_i q_j erfc(a r)/r (:997-1038), the reciprocal
// energy sum Q U / (2 K1 K2 K3) (:1368-1398) and the slab term.  Every
// kernel writes per-block partial sums that the host adds in block order.
constexpr int kEnShort = 10;  // LJ, wall, FENE, kinetic, sum qz, sum q^2, field, W_lj, W_fene, 0

template <typename R>
__device__ __forceinline__ void block_partials(double (&v)[kEnShort], int nt, double *out) {
  __shared__ double red[kEnShort][kBlock / 64];
  for (int t = 0; t < nt; t++) {
    const double x = wave_sum(v[t]);
    if ((threadIdx.x & 63) == 0) red[t][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < nt) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; w++) s += red[threadIdx.x][w];
    out[(size_t)blockIdx.x * nt + threadIdx.x] = s;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_energy_short(const Rep *__restrict__ reps, double *part) {
  const Rep &S = reps[0];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[kEnShort] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < S.NN) {
    const double4 p = S.pos[i];
    // WCA over the full list, half per entry
    const R hLx = (R)S.hLx, hLy = (R)S.hLy, Lx = (R)S.Lx, Ly = (R)S.Ly;
    R xi, yi, zi, qa;
    if constexpr (sizeof(R) == 4) {
      const float4 q = S.pos32[i];
      xi = q.x; yi = q.y; zi = q.z; qa = fabsf(q.w);
    } else {
      xi = p.x; yi = p.y; zi = p.z; qa = fabs(p.w);
    }
    const int n = min(S.nbn1[i], S.nbcap1);
    for (int k = 0; k < n; k++) {
      const int j = S.nb1[(size_t)k * S.NN + i];
      R xj, yj, zj, qj;
      if constexpr (sizeof(R) == 4) {
        const float4 q = S.pos32[j];
        xj = q.x; yj = q.y; zj = q.z; qj = q.w;
      } else {
        const double4 q = S.pos[j];
        xj = q.x; yj = q.y; zj = q.z; qj = q.w;
      }
      R dx = image(xi - xj, Lx, hLx), dy = image(yi - yj, Ly, hLy), dz = zi - zj;
      const R rsq = dx * dx + dy * dy + dz * dz;
      const int m = (int)(qa + fabs(qj));
      R cut, s2;
      if (m == 6) { cut = (R)S.rcl3; s2 = (R)S.sigma2; }
      else if (m < 3) { cut = (R)S.rcl2; s2 = (R)1; }
      else { cut = (R)S.rcl4; s2 = (R)S.sigma3; }
      if (rsq < cut) {
        const R ir2 = s2 / rsq, ir6 = ir2 * ir2 * ir2;
        v[0] += 0.5 * (double)((R)4 * (ir6 * ir6 - ir6) + (R)1);
        v[7] += 0.5 * (double)((R)48 * ir6 * (ir6 - (R)0.5));
      }
    }
    // wall (FP64, as the force)
    const double z = p.z, Lz = S.Lz;
    if (!(fabs(z) < 0.01 || fabs(z - Lz) < 0.01)) {
      const bool big = p.w == 3;
      const double ss = big ? S.sigma_s : 1.0, rw = big ? S.wall_rs : S.wall_r;
      double dl = -1;
      if (z < rw) dl = z;
      else if (Lz - z < rw) dl = Lz - z;
      if (dl > 0) {
        const double t = ss / dl, t3 = t * t * t;
        v[1] = S.wall_c / 3 * (t3 * t3 * t3 - t3);
      }
    }
    // FENE over the bonded partners, half per bond end
    for (int b = S.bstart[i]; b < S.bstart[i + 1]; b++) {
      double dxd, dyd, dzd;
      const double rs = rsq_image(p, S.pos[S.bpart[b]], S.Lx, S.Ly, S.hLx, S.hLy, dxd, dyd, dzd);
      const R R02 = (R)S.R0_2, kR = (R)S.kR02, r2 = (R)rs;
      v[2] += 0.5 * (double)(-(R)0.5 * kR * log((R)1 - r2 / R02));
      v[8] += 0.5 * (double)(-kR / (R02 - r2) * r2);
    }
    const double vx = S.vx[i], vy = S.vy[i], vz = S.vz[i];
    v[3] = (vx * vx + vy * vy + vz * vz) / 2;
    if (S.flags[i] & 2) {
      v[4] = p.w * p.z;
      v[5] = p.w * p.w;
    }
    v[6] = -S.EF * p.w * p.z;
  }
  block_partials<R>(v, kEnShort, part);
}

// Screened real-space pair energy without lB/Beta, added with its pair
// virial r.F into (e, w), for r < r_c; FP32 with the erfc of the FP32 force.
template <typename R>
__device__ __forceinline__ void real_pair_energy(R dx, R dy, R dz, R qq, R rcc2, R alpha, R alpha2, R c2,
                                                 double f, double &e, double &w) {
  const R rsq = dx * dx + dy * dy + dz * dz;
  if (!(rsq < rcc2)) return;
  R ec;
  const R rr = sqrt(rsq);
  if constexpr (sizeof(R) == 4) {
    const float t = 1.0f / (1.0f + 0.3275911f * alpha * rr);
    const float poly =
        t * (0.254829592f +
             t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    ec = poly * expf(-alpha2 * rsq);
  } else {
    ec = erfc(alpha * rr);
  }
  e += f * (double)(qq * ec / rr);
  w += f * (double)(qq * (ec / rr + c2 * exp(-alpha2 * rsq)));
}

// One thread per charge slot: the sweep cells of the last Coulomb evaluation
// (every other charge within r_c once, half per pair), or the walked
// reference-form list (half per entry, repeats included).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_energy_real(const Rep *__restrict__ reps, double *part) {
  using V = typename std::conditional<sizeof(R) == 4, float4, double4>::type;
  const Rep &S = reps[0];
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double v[kEnShort] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (t < S.Nq) {
    const int i = S.charge[t];
    const V *pv = (const V *)(sizeof(R) == 4 ? (const void *)S.pos32 : (const void *)S.pos);
    const V pi = pv[i];
    const R Lx = (R)S.Lx, Ly = (R)S.Ly, hLx = (R)S.hLx, hLy = (R)S.hLy, rcc2 = (R)S.rcc2,
            alpha = (R)S.alpha, alpha2 = (R)S.alpha2, c2 = (R)S.c2;
    auto pair = [&](const V pj) {
      const R dx = image((R)pi.x - (R)pj.x, Lx, hLx), dy = image((R)pi.y - (R)pj.y, Ly, hLy),
              dz = (R)pi.z - (R)pj.z;
      real_pair_energy<R>(dx, dy, dz, (R)pi.w * (R)pj.w, rcc2, alpha, alpha2, c2, 0.5, v[0], v[1]);
    };
    if (S.real_list) {
      const int n = min(S.nbn2[t], S.nbcap2);
      for (int k = 0; k < n; k++) pair(pv[S.nb2[(size_t)k * S.Nq + t]]);
    } else if (S.real_cells) {
      int a, b, c;
      if (cell_of(S.pos[i], S.nc3, S.rcel3, S.hLx, S.hLy, a, b, c)) {
        for (int l = 0; l < 27; l++) {
          const int cl = stencil_cell<true>(l, a, b, c, S.nc3);
          if (cl < 0) continue;
          const int cn = min(S.scnt3[cl], S.ccap3);
          const int *mm = S.mem3 + (size_t)cl * S.ccap3;
          const V *ps = (const V *)S.pos3s + (size_t)cl * S.ccap3;
          for (int k = 0; k < cn; k++)
            if (mm[k] != i) pair(ps[k]);  // the cells hold particle indices
        }
      }
    }
  }
  block_partials<R>(v, 2, part);
}

// sum Q U over the mesh of the last Coulomb evaluation (charge mesh and the
// potential mesh of its inverse FFT).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_energy_mesh(const Rep *__restrict__ reps, double *part) {
  const Rep &S = reps[0];
  const size_t M = (size_t)S.K[0] * S.K[1] * S.K[2];
  const R *q = (const R *)S.meshr, *u = (const R *)S.meshu;
  double v[kEnShort] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x; g < M; g += (size_t)gridDim.x * kBlock)
    v[0] += (double)q[g] * (double)u[g];
  block_partials<R>(v, 1, part);
}

// --------------------------------------------------------- launchers

// One replica (P.nrep == 1); Npe is read from the descriptor.
void launch_observe(const Plan &P, hipStream_t st, int *hist, double *hsum) {
  hipLaunchKernelGGL(k_hist_clear, dim3(1), dim3(kBlock), 0, st, hist);
  const int nb = P.part.total;  // enough blocks for Npe <= NN
  hipLaunchKernelGGL(k_observe, dim3(nb), dim3(kBlock), 0, st, P.reps, hist, hsum);
}

void launch_heights(const Plan &P, hipStream_t st, double *mol) {
  hipLaunchKernelGGL(k_heights, dim3(1), dim3(kBlock), 0, st, P.reps, mol);
}

// With ev != null, records ev[k] before kernel k and ev[8] after the last
// (kernel timing); never used under graph capture.
// With store false the forces are left to the next short-range launch
// (SR_COUL), which assembles and holds them.
template <typename R>
void launch_coulomb(const Plan &P, hipStream_t st, double *out, hipEvent_t *ev, bool store) {
  int k = 0;
  // plane spreading stores the charge mesh only for an evaluation whose
  // forces go out (the energies' mesh term reads it); the held-force
  // evaluation of the step loop keeps the mesh in LDS alone
  const int store_q = out != nullptr;
  auto mark = [&]() {
    if (ev) (void)hipEventRecord(ev[k++], st);
  };
  // the real-space sweep forks onto the plan's second stream once the cells
  // are sorted and joins before the store (timed launches stay in order)
  const bool fork = P.st2 && !ev && P.cell3.total;
  mark();
  hipLaunchKernelGGL(k_bin_q, dim3(P.chg.total), dim3(kBlock), 0, st, P.reps, P.chg, P.pspread ? 1 : 0);
  mark();
  if (P.cell3.total)
    // many cells (batches): one wavefront per cell, whose bitonic stages'
    // barriers stay inside it (306 -> 112 us for the 256-replica stream);
    // few cells: four (11 against 14 us at the default brush)
    hipLaunchKernelGGL(k_sort_q<R>, dim3(P.cell3.total), dim3(P.cell3.total > 4096 ? 64 : kBlock),
                       P.lds_sort3, st, P.reps, P.cell3);
  mark();
  // the sweep's workgroups ride along with the spreading and gather launches
  // (half of their items in each), unless its launches are timed or forked
  const bool ride = P.ride && !ev && !fork && P.cell3.total;
  const int nr = ride ? P.cell3c.total : 0;
  if (fork) {
    (void)hipEventRecord(P.ev_fork, st);
    (void)hipStreamWaitEvent(P.st2, P.ev_fork, 0);
    if (P.real_wg)
      hipLaunchKernelGGL((k_real_cells<R, true>), dim3(P.cell3c.total), dim3(kBlock), 0, P.st2, P.reps,
                         P.cell3c);
    else
      hipLaunchKernelGGL((k_real_cells<R, false>), dim3(P.cell3c.total), dim3(kBlock), 0, P.st2, P.reps,
                         P.cell3c);
    (void)hipEventRecord(P.ev_join, P.st2);
  }
  // plane spreading: the spreading is the forward plane kernel's first
  // part, and the whole sweep rides in the gather launch
  if (P.pspread) {
  } else if (ride && P.real_wg)
    hipLaunchKernelGGL((k_spread_real<R, true>), dim3(P.brick.total + nr), dim3(kBlock), 0, st, P.reps,
                       P.brick, P.cell3c);
  else if (ride)
    hipLaunchKernelGGL((k_spread_real<R, false>), dim3(P.brick.total + nr), dim3(kBlock), 0, st, P.reps,
                       P.brick, P.cell3c);
  else
    hipLaunchKernelGGL(k_spread<R>, dim3(P.brick.total), dim3(kBlock), 0, st, P.reps, P.brick);
  mark();
  // replicas with fused planes take k_plane_*, the others the x and y passes;
  // with plane spreading the workgroup-form sweep's items are dealt to
  // P.ride_parts launches: the gather's alone (1), the plane and the gather
  // launches (2), or the plane, z-pass and gather launches (3)
  const int rparts = ride && P.pspread && P.real_wg ? P.ride_parts : (ride && P.pspread ? 1 : 0);
  const size_t real512 = RealWG<R, kPlaneBlock>::bytes;
  if (P.plane.total && rparts >= 2)
    hipLaunchKernelGGL(k_plane_spread_real<R>, dim3(P.plane.total + nr), dim3(kPlaneBlock),
                       P.lds_plane > real512 ? P.lds_plane : real512, st, P.reps, P.plane, P.cell3c,
                       rparts, store_q, std::min(P.ride_lead, nr));
  else if (P.plane.total && P.pspread)
    hipLaunchKernelGGL(k_plane_spread<R>, dim3(P.plane.total), dim3(kPlaneBlock), P.lds_plane, st,
                       P.reps, P.plane, store_q);
  else if (P.plane.total)
    hipLaunchKernelGGL(k_plane_fwd<R>, dim3(P.plane.total), dim3(kPlaneBlock), P.lds_plane, st, P.reps,
                       P.plane);
  if (P.fx.total)
    hipLaunchKernelGGL(k_fftx_fwd<R>, dim3(P.fx.total), dim3(kBlock), P.lds_x, st, P.reps, P.fx);
  mark();
  if (P.fy.total)
    hipLaunchKernelGGL((k_ffty<R, -1>), dim3(P.fy.total), dim3(kBlock), P.lds_y, st, P.reps, P.fy);
  mark();
  // 512 lanes: the strided line loads and stores spread over more wavefronts
  if (rparts == 3)
    hipLaunchKernelGGL(k_fftz_real<R>, dim3(P.fz.total + nr), dim3(kPlaneBlock),
                       P.lds_z > real512 ? P.lds_z : real512, st, P.reps, P.fz, P.cell3c, rparts,
                       std::min(P.ride_lead, nr));
  else
    hipLaunchKernelGGL(k_fftz_conv<R>, dim3(P.fz.total), dim3(kPlaneBlock), P.lds_z, st, P.reps, P.fz);
  mark();
  if (P.fy.total)
    hipLaunchKernelGGL((k_ffty<R, +1>), dim3(P.fy.total), dim3(kBlock), P.lds_y, st, P.reps, P.fy);
  mark();
  if (P.plane.total)
    hipLaunchKernelGGL(k_plane_inv<R>, dim3(P.plane.total), dim3(kPlaneBlock), P.lds_plane, st, P.reps,
                       P.plane);
  if (P.fx.total)
    hipLaunchKernelGGL(k_fftx_inv<R>, dim3(P.fx.total), dim3(kBlock), P.lds_x, st, P.reps, P.fx);
  mark();
  const int nparts = P.pspread ? rparts : 2;  // the gather launch takes the last part
  if (ride && P.real_wg)
    hipLaunchKernelGGL((k_gather_real<R, true>), dim3(P.chgG.total + nr), dim3(kBlock), 0, st, P.reps,
                       P.chgG, P.cell3c, nparts, std::min(P.ride_lead_g, nr));
  else if (ride)
    hipLaunchKernelGGL((k_gather_real<R, false>), dim3(P.chgG.total + nr), dim3(kBlock), 0, st, P.reps,
                       P.chgG, P.cell3c, nparts, std::min(P.ride_lead_g, nr));
  else
    hipLaunchKernelGGL(k_gather<R>, dim3(P.chgG.total), dim3(kBlock), 0, st, P.reps, P.chgG);
  mark();
  if (fork) (void)hipStreamWaitEvent(st, P.ev_join, 0);
  else if (P.cell3.total && !ride) {
    if (P.real_wg)
      hipLaunchKernelGGL((k_real_cells<R, true>), dim3(P.cell3c.total), dim3(kBlock), 0, st, P.reps,
                         P.cell3c);
    else
      hipLaunchKernelGGL((k_real_cells<R, false>), dim3(P.cell3c.total), dim3(kBlock), 0, st, P.reps,
                         P.cell3c);
  }
  if (P.rlst.total)  // replicas whose real-space force walks the reference-form list
    hipLaunchKernelGGL(k_real_list<R>, dim3(P.rlst.total), dim3(kBlock), 0, st, P.reps, P.rlst);
  if (store || out)
    hipLaunchKernelGGL(k_coul_store<R>, dim3(P.chg.total), dim3(kBlock), 0, st, P.reps, P.chg, out);
  mark();
}

template <typename R>
void launch_short(const Plan &P, hipStream_t st, int mode, int step, const double *eta,
                  double *out) {
  // the Coulomb step's instantiation for every step of a one-replica engine
  // (the default brush: 64.8k against 64.3k steps/s with the plain one on
  // the other steps); batches take the plain one on the steps without the
  // Coulomb evaluation, whose registers it keeps lower (the 256-replica
  // batch: 375k against 358k replica-steps/s)
  if ((mode & SR_COUL) || P.nrep == 1)
    hipLaunchKernelGGL((k_short_range<R, true>), dim3(P.partG.total), dim3(kSRBlock), 0, st, P.reps,
                       P.partG, mode, step, eta, out, P.lq, P.lq_cap, P.lq_nx);
  else
    hipLaunchKernelGGL((k_short_range<R, false>), dim3(P.partG.total), dim3(kSRBlock), 0, st, P.reps,
                       P.partG, mode, step, eta, out, P.lq, P.lq_cap, P.lq_nx);
}

void launch_kick_drift(const Plan &P, hipStream_t st, int store_v) {
  hipLaunchKernelGGL(k_kick_drift, dim3(P.part.total), dim3(kBlock), 0, st, P.reps, P.part, P.lq,
                     store_v);
}

// update_verlet_list's LJ rebuild (compt_acc.f90:717-720); every kernel
// returns at once for replicas whose rebuild flag is clear.
void launch_build(const Plan &P, hipStream_t st, int inc_step) {
  hipLaunchKernelGGL(k_build_lj, dim3(P.grid_build), dim3(kBlock), 0, st, P.reps, P.nrep, P.lq,
                     P.lq_cap, inc_step);
}

void launch_lists(const Plan &P, hipStream_t st, int inc_step, hipEvent_t *ev, bool clear_queue) {
  if (ev) (void)hipEventRecord(ev[0], st);
  hipLaunchKernelGGL(k_bin_lj, dim3(P.grid_bin), dim3(kBlock), 0, st, P.reps, P.lq, P.lq_cap);
  if (ev) (void)hipEventRecord(ev[1], st);
  if (ev) (void)hipEventRecord(ev[2], st);
  hipLaunchKernelGGL(k_build_lj, dim3(P.grid_build), dim3(kBlock), 0, st, P.reps, P.nrep, P.lq,
                     P.lq_cap, inc_step);
  if (ev) (void)hipEventRecord(ev[3], st);
  if (clear_queue) hipLaunchKernelGGL(k_lq_clear, dim3(1), dim3(1), 0, st, P.lq);
}

// The reference-form real-space list (real_verlet_list): every replica
// that keeps one (all), or the replicas whose real-space force walks it.
void launch_real_list(const Plan &P, hipStream_t st, bool all) {
  const WMap &q = all ? P.chg : P.rlst, &c = all ? P.cell2 : P.cell2s;
  if (!c.total) return;
  hipLaunchKernelGGL(k_bin_rl, dim3(q.total), dim3(kBlock), 0, st, P.reps, q);
  hipLaunchKernelGGL(k_sort_rl, dim3(c.total), dim3(kBlock), P.lds_sort, st, P.reps, c);
  hipLaunchKernelGGL(k_build_rl, dim3(q.total), dim3(kBlock), 0, st, P.reps, q);
}

void launch_set_flags(const Plan &P, hipStream_t st, int clear) {
  hipLaunchKernelGGL(k_set_flags, dim3(1), dim3(kBlock), 0, st, P.reps, P.nrep, clear, P.lq,
                     P.lq_cap);
}

void launch_rescale(const Plan &P, hipStream_t st, int gated, int redrift) {
  hipLaunchKernelGGL(k_vsum, dim3(P.part.total), dim3(kBlock), 0, st, P.reps, P.part, P.vpart, 0,
                     gated);
  hipLaunchKernelGGL(k_vsum, dim3(P.part.total), dim3(kBlock), 0, st, P.reps, P.part, P.vpart, 1,
                     gated);
  hipLaunchKernelGGL(k_vscale, dim3(P.part.total), dim3(kBlock), 0, st, P.reps, P.part, P.vpart,
                     gated, redrift);
}

// Energies of one replica (P.nrep == 1) into per-block partials: nb_short
// blocks of kEnShort sums, then nb_q blocks of 2 (real space), then nb_mesh
// blocks of 1 (mesh); the Coulomb evaluation must have run at the current
// positions (cells sorted, meshes filled).
template <typename R>
void launch_energies(const Plan &P, hipStream_t st, double *part, int nb_short, int nb_q, int nb_mesh) {
  hipLaunchKernelGGL(k_energy_short<R>, dim3(nb_short), dim3(kBlock), 0, st, P.reps, part);
  double *p2 = part + (size_t)nb_short * kEnShort;
  if (nb_q) hipLaunchKernelGGL(k_energy_real<R>, dim3(nb_q), dim3(kBlock), 0, st, P.reps, p2);
  if (nb_mesh)
    hipLaunchKernelGGL(k_energy_mesh<R>, dim3(nb_mesh), dim3(kBlock), 0, st, P.reps, p2 + (size_t)nb_q * 2);
}
template void launch_energies<float>(const Plan &, hipStream_t, double *, int, int, int);
template void launch_energies<double>(const Plan &, hipStream_t, double *, int, int, int);

template void launch_coulomb<float>(const Plan &, hipStream_t, double *, hipEvent_t *, bool);
template void launch_coulomb<double>(const Plan &, hipStream_t, double *, hipEvent_t *, bool);
template void launch_short<float>(const Plan &, hipStream_t, int, int, const double *, double *);
template void launch_short<double>(const Plan &, hipStream_t, int, int, const double *, double *);

}  // namespace mdk
