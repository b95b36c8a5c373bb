// md_ewald.hip -- the standard-Ewald reference force of error_analysis
// (compt_acc.f90:729-812) on the device, FP64: a spot check of the SPME
// path's accuracy for a replica, computed from its resident positions.
//
// Standard_Ewald (compt_acc.f90:1052-1135) sums, over the k-vectors of the
// ellipsoid (p/K1)^2 + (q/K2)^2 + (r/K3)^2 <= 1 in the half space r >= 0,
// S(k) = sum_m q_m exp(i k.r_m) and F_m += e_k Im(q_m exp(i k.r_m) conj(S(k))) k.
// The reference builds exp(i k.r_m) from per-axis power tables; here each
// phase is one sincos of k.r_m (the same value to rounding).  Two dense
// passes: k-vectors x charges for S (a thread per k-vector, charges staged
// through LDS), then charges x k-vectors for F (a thread per charge,
// k-vectors and S staged through LDS).  real_space at the check's
// parameters (compt_acc.f90:998-1010, every pair within r_c with the x,y
// minimum image) runs as an all-pairs sweep, a thread per target charge.
// Each pass splits its inner range into slices whose partial sums are added
// in slice order, so the result does not depend on scheduling.
#include <mutex>

#include "md_kernels.h"

namespace mdk {

namespace {

constexpr int kEB = 256;  // threads per workgroup and LDS stage of every pass

// Device scratch of the allocation-free error analysis: every GPU holds its
// own copy from library load to process end.  In host code its name is not a
// device address; ewald_scratch() returns the address on the current device.
__device__ __align__(256) char scratch[kEwaldScratchBytes];

__device__ __forceinline__ int slice_begin(int n, int j, int ns) {
  return (int)((long long)n * j / ns);
}

// charges packed as (x, y, z, q)
__global__ __launch_bounds__(kEB) void k_ew_pack(EwaldArgs a) {
  const int m = blockIdx.x * kEB + threadIdx.x;
  if (m < a.Nq) a.cq[m] = a.pos[a.charge[m]];
}

// Partial structure factors of slice blockIdx.y of the charges.
__global__ __launch_bounds__(kEB) void k_ew_sfac(EwaldArgs a) {
  __shared__ double4 c[kEB];
  const int k = blockIdx.x * kEB + threadIdx.x;
  const int m0 = slice_begin(a.Nq, blockIdx.y, a.ms), m1 = slice_begin(a.Nq, blockIdx.y + 1, a.ms);
  const double4 kk = k < a.nk ? a.kv[k] : make_double4(0, 0, 0, 0);
  double sr = 0, si = 0;
  for (int base = m0; base < m1; base += kEB) {
    const int n = min(kEB, m1 - base);
    __syncthreads();
    if (threadIdx.x < n) c[threadIdx.x] = a.cq[base + threadIdx.x];
    __syncthreads();
    for (int j = 0; j < n; j++) {
      const double4 q = c[j];
      double sn, cs;
      sincos(kk.x * q.x + kk.y * q.y + kk.z * q.z, &sn, &cs);
      sr += q.w * cs;
      si += q.w * sn;
    }
  }
  if (k < a.nk) a.sp[(size_t)blockIdx.y * a.nk + k] = make_double2(sr, si);
}

__global__ __launch_bounds__(kEB) void k_ew_ssum(EwaldArgs a) {
  const int k = blockIdx.x * kEB + threadIdx.x;
  if (k >= a.nk) return;
  double sr = 0, si = 0;
  for (int j = 0; j < a.ms; j++) {
    const double2 v = a.sp[(size_t)j * a.nk + k];
    sr += v.x;
    si += v.y;
  }
  a.sk[k] = make_double2(sr, si);
}

// Reciprocal force on each charge from slice blockIdx.y of the k-vectors.
__global__ __launch_bounds__(kEB) void k_ew_force(EwaldArgs a) {
  __shared__ double4 kvs[kEB];
  __shared__ double2 ss[kEB];
  const int m = blockIdx.x * kEB + threadIdx.x;
  const int k0 = slice_begin(a.nk, blockIdx.y, a.ks), k1 = slice_begin(a.nk, blockIdx.y + 1, a.ks);
  const double4 c = m < a.Nq ? a.cq[m] : make_double4(0, 0, 0, 0);
  double fx = 0, fy = 0, fz = 0;
  for (int base = k0; base < k1; base += kEB) {
    const int n = min(kEB, k1 - base);
    __syncthreads();
    if (threadIdx.x < n) {
      kvs[threadIdx.x] = a.kv[base + threadIdx.x];
      ss[threadIdx.x] = a.sk[base + threadIdx.x];
    }
    __syncthreads();
    for (int j = 0; j < n; j++) {
      const double4 kk = kvs[j];
      const double2 S = ss[j];
      double sn, cs;
      sincos(kk.x * c.x + kk.y * c.y + kk.z * c.z, &sn, &cs);
      // Im(q e^{ik.r} conj(S)) e_k
      const double w = c.w * (sn * S.x - cs * S.y) * kk.w;
      fx += w * kk.x;
      fy += w * kk.y;
      fz += w * kk.z;
    }
  }
  if (m < a.Nq) {
    double *o = a.rp + (size_t)blockIdx.y * 3 * a.Nq;
    o[m] = fx;
    o[(size_t)a.Nq + m] = fy;
    o[2 * (size_t)a.Nq + m] = fz;
  }
}

// real_space pair force on each target charge from slice blockIdx.y of the
// source charges: every other charge within r_c, x,y minimum image.
__global__ __launch_bounds__(kEB) void k_ew_real(EwaldArgs a) {
  __shared__ double4 src[kEB];
  const int m = blockIdx.x * kEB + threadIdx.x;
  const int s0 = slice_begin(a.Nq, blockIdx.y, a.rs), s1 = slice_begin(a.Nq, blockIdx.y + 1, a.rs);
  const double4 c = m < a.Nq ? a.cq[m] : make_double4(0, 0, 0, 0);
  double fx = 0, fy = 0, fz = 0;
  for (int base = s0; base < s1; base += kEB) {
    const int n = min(kEB, s1 - base);
    __syncthreads();
    if (threadIdx.x < n) src[threadIdx.x] = a.cq[base + threadIdx.x];
    __syncthreads();
    for (int j = 0; j < n; j++) {
      if (base + j == m) continue;
      const double4 q = src[j];
      double dx = c.x - q.x, dy = c.y - q.y;
      const double dz = c.z - q.z;
      if (dx > a.hLx) dx -= a.Lx;
      else if (dx <= -a.hLx) dx += a.Lx;
      if (dy > a.hLy) dy -= a.Ly;
      else if (dy <= -a.hLy) dy += a.Ly;
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < a.rcc2) {
        const double rr = sqrt(rsq);
        const double ff =
            c.w * q.w / rsq * (erfc(a.alpha * rr) / rr + a.c2 * exp(-a.alpha2 * rsq)) * a.c1;
        fx += ff * dx;
        fy += ff * dy;
        fz += ff * dz;
      }
    }
  }
  if (m < a.Nq) {
    double *o = a.qp + (size_t)blockIdx.y * 3 * a.Nq;
    o[m] = fx;
    o[(size_t)a.Nq + m] = fy;
    o[2 * (size_t)a.Nq + m] = fz;
  }
}

// M_z = sum q z lB/Beta (compt_acc.f90:1040-1043): one workgroup, fixed order.
__global__ __launch_bounds__(kEB) void k_ew_mz(EwaldArgs a) {
  __shared__ double red[kEB];
  double s = 0;
  for (int m = threadIdx.x; m < a.Nq; m += kEB) s += a.cq[m].w * a.cq[m].z * a.c1;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kEB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *a.mz = red[0];
}

// Slices added in order, slab correction (compt_acc.f90:1044-1047), written
// at the particle index into out(NN, 3).
__global__ __launch_bounds__(kEB) void k_ew_final(EwaldArgs a) {
  const int m = blockIdx.x * kEB + threadIdx.x;
  if (m >= a.Nq) return;
  double f[3];
  for (int d = 0; d < 3; d++) {
    double s = 0;
    for (int j = 0; j < a.rs; j++) s += a.qp[((size_t)j * 3 + d) * a.Nq + m];
    double t = 0;
    for (int j = 0; j < a.ks; j++) t += a.rp[((size_t)j * 3 + d) * a.Nq + m];
    f[d] = s + t;
  }
  f[2] -= a.slab * a.cq[m].w * *a.mz;
  const int i = a.charge[m];
  a.out[i] = f[0];
  a.out[(size_t)a.NN + i] = f[1];
  a.out[2 * (size_t)a.NN + i] = f[2];
}

// The k-vectors of Standard_Ewald as the whole box r = 0..K3, p = -K1..K1,
// q = -K2..K2 in the reference's loop order (compt_acc.f90:650-668), e_k = 0
// outside the ellipsoid (kcut > 1, in default real) and at k = 0, so the
// passes need no compaction (a vector of weight 0 adds nothing).
__global__ __launch_bounds__(kEB) void k_ew_kbox(int nk, double4 *kv, int K1, int K2, int K3, double c1,
                                                 double c2, double c3, double pref, double alpha2) {
  const int k = blockIdx.x * kEB + threadIdx.x;
  if (k >= nk) return;
  const int w2 = 2 * K2 + 1, w12 = (2 * K1 + 1) * w2;
  const int r = k / w12, pp = (k - r * w12) / w2 - K1, q = k % w2 - K2;
  const float fa = 1.0f * pp / K1, fb = 1.0f * q / K2, fc = 1.0f * r / K3;
  const float kcut = fa * fa + fb * fb + fc * fc;
  const double kx = pp * c1, ky = q * c2, kz = r * c3;
  const double ksqr = kx * kx + ky * ky + kz * kz;
  double ek = 0;
  if (!(kcut > 1 || kcut == 0)) ek = (r == 0 ? 1 : 2) * pref * exp(-ksqr / 4 / alpha2) / ksqr;
  kv[k] = make_double4(kx, ky, kz, ek);
}

__global__ __launch_bounds__(kEB) void k_ew_zero(double *x, size_t n) {
  for (size_t k = (size_t)blockIdx.x * kEB + threadIdx.x; k < n; k += (size_t)gridDim.x * kEB) x[k] = 0;
}

// rmsf sums over every particle and sum q^2 over the charges, one
// workgroup in a fixed order: out[0] = sum |F_se - F_c|^2, out[1] =
// sum |F_se|^2, out[2] = sum q^2.
__global__ __launch_bounds__(kEB) void k_ew_stats(const double *fse, const double *fc, size_t n3,
                                                  EwaldArgs a, double *out) {
  __shared__ double red[3][kEB];
  double s1 = 0, s2 = 0, s3 = 0;
  for (size_t k = threadIdx.x; k < n3; k += kEB) {
    const double d = fse[k] - fc[k];
    s1 += d * d;
    s2 += fse[k] * fse[k];
  }
  for (int m = threadIdx.x; m < a.Nq; m += kEB) s3 += a.cq[m].w * a.cq[m].w;
  red[0][threadIdx.x] = s1;
  red[1][threadIdx.x] = s2;
  red[2][threadIdx.x] = s3;
  __syncthreads();
  for (int w = kEB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int t = 0; t < 3; t++) red[t][threadIdx.x] += red[t][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

void *ewald_scratch() {
  // the address of `scratch` on each device, taken once per device with that
  // device current (fixed table: no allocation)
  constexpr int kMaxDev = 64;
  static std::mutex mu;
  static void *ptr[kMaxDev] = {nullptr};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    ptr[dev] = p;
  }
  return ptr[dev];
}

void launch_ewald_box(int nk, double4 *kv, hipStream_t st, int K1, int K2, int K3, double c1, double c2,
                      double c3, double pref, double alpha2) {
  hipLaunchKernelGGL(k_ew_kbox, dim3((nk + kEB - 1) / kEB), dim3(kEB), 0, st, nk, kv, K1, K2, K3, c1, c2,
                     c3, pref, alpha2);
}

void launch_zero(double *x, size_t n, hipStream_t st) {
  const int nb = (int)std::min<size_t>(1024, (n + kEB - 1) / kEB);
  hipLaunchKernelGGL(k_ew_zero, dim3(nb > 0 ? nb : 1), dim3(kEB), 0, st, x, n);
}

void launch_ewald_stats(const double *fse, const double *fc, size_t n3, const EwaldArgs &a, double *out,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_ew_stats, dim3(1), dim3(kEB), 0, st, fse, fc, n3, a, out);
}

void launch_ewald(const EwaldArgs &a, hipStream_t st) {
  const int bq = (a.Nq + kEB - 1) / kEB, bk = (a.nk + kEB - 1) / kEB;
  hipLaunchKernelGGL(k_ew_pack, dim3(bq), dim3(kEB), 0, st, a);
  hipLaunchKernelGGL(k_ew_sfac, dim3(bk, a.ms), dim3(kEB), 0, st, a);
  hipLaunchKernelGGL(k_ew_ssum, dim3(bk), dim3(kEB), 0, st, a);
  hipLaunchKernelGGL(k_ew_force, dim3(bq, a.ks), dim3(kEB), 0, st, a);
  hipLaunchKernelGGL(k_ew_real, dim3(bq, a.rs), dim3(kEB), 0, st, a);
  hipLaunchKernelGGL(k_ew_mz, dim3(1), dim3(kEB), 0, st, a);
  hipLaunchKernelGGL(k_ew_final, dim3(bq), dim3(kEB), 0, st, a);
}

}  // namespace mdk
