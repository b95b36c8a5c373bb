// md_kernels.h -- launch plan and kernel launchers of the brush engine.
#pragma once
#include "../../include/md_engine.h"
#include "md_device.h"

namespace mdk {

// Mode bits of the short-range kernel.
enum {
  SR_HELD = 1,           // start from the held Coulomb force acc_c
  SR_NOISE_ETA = 2,      // add injected normals eta(NN,3) * sigma
  SR_NOISE_PHILOX = 4,   // add Philox normals * sigma
  SR_FIELD = 8,          // add q * E_z
  SR_ANCHOR = 16,        // zero the force on anchors
  SR_KICK = 32,          // second half kick of new_position
  SR_DRMAX = 64,         // dr_max bookkeeping and rebuild flags
  SR_DRIFT = 128,        // with SR_KICK, for replayed steps: v holds v(n-1) (the
                         // first half kick is redone), and the next step's
                         // kick and drift go to the other position buffer
  SR_BIN = 256,          // with SR_DRMAX: every workgroup decides and bins its
                         // particles for a rebuild (the step's list kernel is
                         // then k_build_lj alone, which does the bookkeeping)
  SR_COUL = 512          // with SR_HELD, after launch_coulomb(store = false):
                         // assemble, hold and add the evaluation's Coulomb force
};

// Complex points in each of the two LDS regions of a fused FFT plane
// kernel: the x lines (K2/2 rows of K1) or the y lines (Kh lines of K2,
// laid out K2 + 1 apart so that the transposing accesses, which step from
// line to line, fall in different LDS banks).
__host__ __device__ inline int plane_line_stride(int K2) { return K2 + 1; }
// ... and of the z pass's lines (K3 points, K3 + 1 apart), for the same reason.
__host__ __device__ inline int zline_stride(int K3) { return K3 + 1; }
// (rows are paired as one complex row; an odd K2 leaves one row alone)
__host__ __device__ inline size_t plane_region(int K1, int K2, int Kh) {
  const size_t a = (size_t)((K2 + 1) / 2) * K1, b = (size_t)Kh * plane_line_stride(K2);
  return a > b ? a : b;
}

// Everything a launch needs: device descriptors and the block maps.
struct Plan {
  Rep *reps = nullptr;
  int nrep = 0;
  // block maps: particles (thread each), charges, charges x G lanes, FFT
  // tiles, LJ cells (workgroup each, wavefront each), reference real-list
  // cells, real-space sweep cells (workgroup each, per target chunk)
  // ... mesh bricks (one wavefront each), z-planes
  WMap part{}, chg{}, chgG{}, fx{}, fy{}, fz{}, cell1{}, cell1w{}, cell2{}, cell3{}, cell3c{},
      brick{}, plane{}, partG{};  // partG: short-range workgroups of kSRParticles
  // charges and reference-form list cells of the replicas whose real-space
  // force walks that list (Rep::real_list)
  WMap rlst{}, cell2s{};
  size_t lds_x = 0, lds_y = 0, lds_z = 0, lds_sort = 0, lds_sort3 = 0, lds_plane = 0;
  // LJ rebuild queue (replicas flagged this step) and the fixed grids of the
  // list kernels, which stride over the queued replicas' particles
  unsigned long long *lq = nullptr;
  unsigned long long *lq_nx = nullptr;  // the queue of the other position buffer's steps
  int lq_cap = 0;
  int grid_bin = 1, grid_build = 1;
  double *vpart = nullptr;  // partial sums of rescale_velocity
  bool fp32 = false;
  bool real_wg = false;     // real-space sweep form of the engine (workgroup per item)
  bool ride = true;         // sweep workgroups inside the spreading and gather launches
  bool pspread = false;     // spreading inside the forward plane kernel (k_plane_spread)
  int ride_parts = 3;       // with plane spreading: launches the sweep's items ride in (1..3)
  int ride_lead = 0;    // riding sweep workgroups dispatched before the plane launch's own
  int ride_lead_g = 0;  // ... before the gather launch's own
  // second stream of a Coulomb evaluation (the real-space sweep runs there,
  // beside the SPME kernels) and its fork / join events; null: one stream
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

template <typename R>
void launch_coulomb(const Plan &P, hipStream_t st, double *out, hipEvent_t *ev = nullptr,
                    bool store = true);
template <typename R>
void launch_short(const Plan &P, hipStream_t st, int mode, int step, const double *eta,
                  double *out);
// store_v = 0: the first step of a run of replayed steps (SR_DRIFT)
void launch_kick_drift(const Plan &P, hipStream_t st, int store_v = 1);
// clear_queue: empty the rebuild queue afterwards (list updates outside the
// replayed step cycle, where the next kick does it)
void launch_build(const Plan &P, hipStream_t st, int inc_step);
void launch_lists(const Plan &P, hipStream_t st, int inc_step, hipEvent_t *ev = nullptr,
                  bool clear_queue = false);
void launch_real_list(const Plan &P, hipStream_t st, bool all);
void launch_set_flags(const Plan &P, hipStream_t st, int clear);
// gated: only batch replicas whose pre-heat rule is due (main.f90:67-75);
// redrift: between replayed steps, whose next drift already ran with the
// velocities before the rescale: redo it with the rescaled ones
void launch_rescale(const Plan &P, hipStream_t st, int gated = 0, int redrift = 0);
void launch_observe(const Plan &P, hipStream_t st, int *hist, double *hsum);
// one replica (P.nrep == 1): energies into per-block partial sums
constexpr int kEnShortTerms = 10;
template <typename R>
void launch_energies(const Plan &P, hipStream_t st, double *part, int nb_short, int nb_q, int nb_mesh);
// one replica (P.nrep == 1): per-molecule heights, 4 doubles per molecule
void launch_heights(const Plan &P, hipStream_t st, double *mol);

// Standard-Ewald reference force of error_analysis (FP64, one replica).
struct EwaldArgs {
  int Nq = 0, NN = 0, nk = 0;
  int ms = 1, ks = 1, rs = 1;       // slices: charges (S), k-vectors (F), sources (real)
  const int *charge = nullptr;      // particle index of charge m
  const double4 *pos = nullptr;     // canonical positions (x, y, z, q)
  double Lx = 0, Ly = 0, hLx = 0, hLy = 0;
  double alpha = 0, alpha2 = 0, c2 = 0, rcc2 = 0;  // c2 = 2 alpha / sqrt(pi)
  double c1 = 0, slab = 0;          // lB/Beta, 4 pi / (Lx Ly Lz Zfac)
  const double4 *kv = nullptr;      // k-vectors (kx, ky, kz, e_k)
  double4 *cq = nullptr;            // charges (x, y, z, q)
  double2 *sp = nullptr, *sk = nullptr;  // partial / summed structure factors
  double *rp = nullptr, *qp = nullptr;   // reciprocal / real-space partial forces
  double *mz = nullptr;
  double *out = nullptr;            // (NN, 3): charged particles written
};
void launch_ewald(const EwaldArgs &a, hipStream_t st);

// The allocation-free form: its buffers are carved from a file-scope device
// scratch array of kEwaldScratchBytes (one per GPU); ewald_scratch() gives
// its address on the current device (null on failure).  k-vectors as the
// whole box with zero weight outside the ellipsoid (launch_ewald_box), a fill
// kernel in place of memsets, and the rmsf sums on the device.
constexpr size_t kEwaldScratchBytes = (size_t)64 << 20;
void *ewald_scratch();
void launch_ewald_box(int nk, double4 *kv, hipStream_t st, int K1, int K2, int K3, double c1, double c2,
                      double c3, double pref, double alpha2);
void launch_zero(double *x, size_t n, hipStream_t st);
void launch_ewald_stats(const double *fse, const double *fc, size_t n3, const EwaldArgs &a, double *out,
                        hipStream_t st);

// Initialize_ewald_parameters (compt_acc.f90:387-414) at (tol, tau_rf):
// alpha, r_c, r_v, the k-ranges before the mesh factor 3, and whether the
// real space keeps a Verlet list; ovr_alpha / ovr_rcc apply when use_ovr.
void ewald_rule(const md_params *p, double tol, double tau_rf, bool use_ovr, double *alpha,
                double *rcc, double *rvc, int k[3], int *real_verlet);

}  // namespace mdk
