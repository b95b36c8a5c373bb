// md_kernels.h -- launch plan and kernel launchers of the brush engine.
#pragma once
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
  SR_DRIFT = 128         // with SR_KICK, for replayed steps: v holds v(n-1) (the
                         // first half kick is redone), and the next step's
                         // kick and drift go to the other position buffer
};

// Everything a launch needs: device descriptors and the block maps.
struct Plan {
  Rep *reps = nullptr;
  int nrep = 0;
  // block maps: particles (thread each), charges, charges x G lanes, FFT
  // tiles, LJ cells (workgroup each, wavefront each), reference real-list
  // cells, real-space sweep cells (workgroup each, per target chunk)
  // ... mesh bricks (one wavefront each), z-planes
  WMap part{}, chg{}, chgG{}, fx{}, fy{}, fz{}, cell1{}, cell1w{}, cell2{}, cell3{}, cell3c{},
      brick{}, plane{}, partG{};  // partG: particles x kSRLanes lanes
  size_t lds_x = 0, lds_y = 0, lds_z = 0, lds_sort = 0, lds_sort3 = 0, lds_plane = 0;
  // LJ rebuild queue (replicas flagged this step) and the fixed grids of the
  // list kernels, which stride over the queued replicas' particles
  unsigned long long *lq = nullptr;
  unsigned long long *lq_nx = nullptr;  // the queue of the other position buffer's steps
  int lq_cap = 0;
  int grid_bin = 1, grid_build = 1;
  double *vpart = nullptr;  // partial sums of rescale_velocity
  bool fp32 = false;
};

template <typename R>
void launch_coulomb(const Plan &P, hipStream_t st, double *out, hipEvent_t *ev = nullptr);
template <typename R>
void launch_short(const Plan &P, hipStream_t st, int mode, int step, const double *eta,
                  double *out);
// store_v = 0: the first step of a run of replayed steps (SR_DRIFT)
void launch_kick_drift(const Plan &P, hipStream_t st, int store_v = 1);
// clear_queue: empty the rebuild queue afterwards (list updates outside the
// replayed step cycle, where the next kick does it)
void launch_lists(const Plan &P, hipStream_t st, int inc_step, hipEvent_t *ev = nullptr,
                  bool clear_queue = false);
void launch_real_list(const Plan &P, hipStream_t st);
void launch_set_flags(const Plan &P, hipStream_t st, int clear);
// gated: only batch replicas whose pre-heat rule is due (main.f90:67-75);
// redrift: between replayed steps, whose next drift already ran with the
// velocities before the rescale: redo it with the rescaled ones
void launch_rescale(const Plan &P, hipStream_t st, int gated = 0, int redrift = 0);
void launch_observe(const Plan &P, hipStream_t st, int *hist, double *hsum);

}  // namespace mdk
