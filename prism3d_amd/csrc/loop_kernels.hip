// The tail of one inner fixed-point iteration on the device: the state
// snapshot, the abundance and temperature damping, and the percentile
// convergence test of PDRSolver3D.run() (solver_3d.py:262-436).
//
// The convergence figures are exact order statistics, not histogram
// estimates.  The per-cell relative changes are recomputed from the state in
// every pass (nothing per cell is stored), their IEEE bit patterns are
// ordered like their values (all are >= 0), and an 8-bit-digit radix select
// finds the ten ranks the figures need in eight passes.  The ranks and the
// interpolation follow numpy's 'linear' percentile (virtual index (n - 1) q,
// the lerp that switches to the upper end at gamma >= 0.5),
// so the figures equal np.percentile on the same values bit for bit.
//
// Elementwise arithmetic is uncontracted, so it rounds as the reference's
// numpy expressions do.
#include <mutex>

#include "p3d_common.h"

namespace p3d {
namespace {

constexpr int kSel = 10;          // selections (order statistics) per call
constexpr int kBins = 256;        // 8-bit digits
constexpr int kPasses = 8;        // 64-bit keys
constexpr int kHistBlock = 256;   // threads per histogram workgroup
constexpr int kHistGrid = 1024;   // grid-stride workgroups of the histogram passes

// Arrays of the test: 0 = |dT|/T, 1 = |dx_H2| / max(x_H2, 0.01),
// 2 = |dx_CO| / max(x_CO, 1e-6) (solver_3d.py:381-400).
// Selections: T p95 (lo, hi), T p99 (lo, hi), T max, H2 p99 (lo, hi), H2 max,
// CO p99 (lo, hi).
__constant__ int kSelArray[kSel] = {0, 0, 0, 0, 0, 1, 1, 1, 2, 2};

// Workspace layout (device, P3D_CONV_WORKSPACE_BYTES).
struct ConvWs {
  unsigned int hist[kSel][kBins];
  unsigned long long prefix[kSel];  // key bits fixed so far
  unsigned long long rank[kSel];    // rank still to find below the prefix
  int active[kSel];
  unsigned long long count[3];      // values passing each mask
  unsigned long long nan_count[3];
  unsigned long long n_hot;         // rel_T > tol
  double gamma[3][2];               // interpolation weights (T p95, T p99, H2 p99 / CO p99)
};
static_assert(sizeof(ConvWs) <= P3D_CONV_WORKSPACE_BYTES, "workspace too small");

__device__ __forceinline__ unsigned long long key_of(double v) {
  return (unsigned long long)__double_as_longlong(v);
}

// The three masked relative changes of cell i; returns the validity mask
// (bit a set when array a holds a value for this cell).
__device__ __forceinline__ int rel_changes(int64_t i, const double *__restrict__ t_old,
                                           const double *__restrict__ t_gas,
                                           const double *__restrict__ h2_old,
                                           const double *__restrict__ x_h2,
                                           const double *__restrict__ co_old,
                                           const double *__restrict__ x_co, double v[3]) {
#pragma clang fp contract(off)
  int m = 0;
  const double to = t_old[i];
  if (to > 10.0) {  // mask_T = old_T > 10
    v[0] = fabs(t_gas[i] - to) / to;
    m |= 1;
  }
  const double ho = h2_old[i];
  if (ho > 1e-6) {  // mask_H2 = old_x_H2 > 1e-6, floor 0.01
    v[1] = fabs(x_h2[i] - ho) / (ho < 0.01 ? 0.01 : ho);
    m |= 2;
  }
  const double co = co_old[i];
  if (co > 1e-8) {  // mask_CO = old_x_CO > 1e-8, floor 1e-6
    v[2] = fabs(x_co[i] - co) / (co < 1e-6 ? 1e-6 : co);
    m |= 4;
  }
  return m;
}

__global__ void __launch_bounds__(64) conv_reset(ConvWs *__restrict__ ws) {
  const int l = threadIdx.x;
  for (int s = 0; s < kSel; ++s)
    for (int b = l; b < kBins; b += 64) ws->hist[s][b] = 0u;
  if (l < kSel) {
    ws->prefix[l] = 0ull;
    ws->rank[l] = 0ull;
    ws->active[l] = 0;
  }
  if (l < 3) {
    ws->count[l] = 0ull;
    ws->nan_count[l] = 0ull;
  }
  if (l == 0) ws->n_hot = 0ull;
}

// One radix pass: histogram digit `pass` (from the most significant byte) of
// every value whose higher digits equal its selection's prefix.  Pass 0 also
// counts the values, the NaNs and the cells with rel_T > tol.
__global__ void __launch_bounds__(kHistBlock)
conv_hist(int64_t n, int pass, double tol, const double *__restrict__ t_old,
          const double *__restrict__ t_gas, const double *__restrict__ h2_old,
          const double *__restrict__ x_h2, const double *__restrict__ co_old,
          const double *__restrict__ x_co, ConvWs *__restrict__ ws) {
  __shared__ unsigned int h[kSel][kBins];
  __shared__ unsigned long long pre[kSel];
  __shared__ int act[kSel];
  __shared__ unsigned long long cnt[7];
  for (int s = 0; s < kSel; ++s) h[s][threadIdx.x] = 0u;
  if (threadIdx.x < kSel) {
    pre[threadIdx.x] = ws->prefix[threadIdx.x];
    act[threadIdx.x] = pass == 0 ? 1 : ws->active[threadIdx.x];
  }
  if (threadIdx.x < 7) cnt[threadIdx.x] = 0ull;
  __syncthreads();
  const int shift = 56 - 8 * pass;
  unsigned int c[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};  // count x3, nan x3, hot
  for (int64_t i = (int64_t)blockIdx.x * kHistBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kHistBlock) {
    double v[3];
    const int m = rel_changes(i, t_old, t_gas, h2_old, x_h2, co_old, x_co, v);
    if (pass == 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (m & (1 << a)) {
          c[a] += 1u;
          c[3 + a] += v[a] != v[a] ? 1u : 0u;
        }
      if ((m & 1) && v[0] > tol) c[6] += 1u;
    }
#pragma unroll
    for (int s = 0; s < kSel; ++s) {
      const int a = kSelArray[s];
      if (!act[s] || !(m & (1 << a))) continue;
      const unsigned long long k = key_of(v[a]);
      if (pass > 0 && (k >> (shift + 8)) != pre[s]) continue;
      atomicAdd(&h[s][(k >> shift) & 0xFF], 1u);
    }
  }
  if (pass == 0) {
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      unsigned int t = c[j];
      for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o);
      if ((threadIdx.x & 63) == 0 && t) atomicAdd(&cnt[j], (unsigned long long)t);
    }
  }
  __syncthreads();
  for (int s = 0; s < kSel; ++s) {
    const unsigned int b = h[s][threadIdx.x];
    if (b) atomicAdd(&ws->hist[s][threadIdx.x], b);
  }
  if (pass == 0 && threadIdx.x < 7 && cnt[threadIdx.x]) {
    unsigned long long *dst = threadIdx.x < 3   ? &ws->count[threadIdx.x]
                              : threadIdx.x < 6 ? &ws->nan_count[threadIdx.x - 3]
                                                : &ws->n_hot;
    atomicAdd(dst, cnt[threadIdx.x]);
  }
}

// numpy's linear-percentile ranks for q among m values (m >= 1): lo and hi
// order statistics and the weight gamma, as numpy's _quantile, _get_indexes
// and _get_gamma compute them (numpy/lib/_function_base_impl.py).
__device__ void np_ranks(unsigned long long m, double q, unsigned long long &lo,
                         unsigned long long &hi, double &gamma) {
#pragma clang fp contract(off)
  const double v = (double)(m - 1) * q;  // (n - 1) * q of numpy's 'linear' method
  const double prev = floor(v);
  if (v >= (double)m - 1.0) {
    lo = hi = m - 1;
    gamma = v - (-1.0);
  } else if (v < 0.0) {
    lo = hi = 0;
    gamma = v - 0.0;
  } else {
    lo = (unsigned long long)prev;
    hi = lo + 1;
    gamma = v - prev;
  }
}

// After pass `pass`: each selection (one wavefront) finds the digit bin that
// holds its rank, fixes that digit in its prefix and clears the histogram.
// Pass 0 first derives the ranks from the counts.
__global__ void __launch_bounds__(64 * kSel) conv_step(int pass, ConvWs *__restrict__ ws) {
  const int s = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int a = kSelArray[s];
  if (pass == 0) {
    const unsigned long long m = ws->count[a];
    unsigned long long r = 0;
    if (m > 0) {
      unsigned long long lo, hi;
      double g;
      // s: 0/1 T p95, 2/3 T p99, 4 T max, 5/6 H2 p99, 7 H2 max, 8/9 CO p99
      const double q = (s < 2) ? 0.95 : 0.99;
      if (s == 4 || s == 7) {
        r = m - 1;
      } else {
        np_ranks(m, q, lo, hi, g);
        const bool upper = (s == 1 || s == 3 || s == 6 || s == 9);
        r = upper ? hi : lo;
        if (l == 0 && !upper) {
          const int gi = s == 0 ? 0 : (s == 2 ? 1 : 2);
          ws->gamma[gi][a == 2 ? 1 : 0] = g;
        }
      }
    }
    if (l == 0) {
      ws->rank[s] = r;
      ws->active[s] = m > 0 ? 1 : 0;
    }
  }
  __syncthreads();
  if (!ws->active[s]) {
    for (int b = l; b < kBins; b += 64) ws->hist[s][b] = 0u;
    return;
  }
  // four bins per lane, lane-ordered: inclusive scan over 256 bins
  unsigned int c4[4], loc = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    c4[j] = ws->hist[s][4 * l + j];
    loc += c4[j];
  }
  unsigned long long incl = loc;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long t = __shfl_up(incl, o);
    if (l >= o) incl += t;
  }
  const unsigned long long excl = incl - loc;
  const unsigned long long r = ws->rank[s];
  // the lane whose four bins span rank r picks the bin
  if (r >= excl && r < incl) {
    unsigned long long below = excl;
    int b = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (r < below + c4[j]) {
        b = j;
        break;
      }
      below += c4[j];
    }
    ws->rank[s] = r - below;
    ws->prefix[s] = (ws->prefix[s] << 8) | (unsigned long long)(4 * l + b);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) ws->hist[s][4 * l + j] = 0u;
}

// numpy's _lerp(a, b, t): a + (b - a) t, or b - (b - a)(1 - t) when t >= 0.5.
__device__ double np_lerp(double a, double b, double t) {
#pragma clang fp contract(off)
  const double d = b - a;
  return t >= 0.5 ? b - d * (1.0 - t) : a + d * t;
}

__global__ void __launch_bounds__(64) conv_final(double tol, const ConvWs *__restrict__ ws,
                                                 double *__restrict__ rec) {
  if (threadIdx.x != 0) return;
  double val[kSel];
  for (int s = 0; s < kSel; ++s) val[s] = __longlong_as_double((long long)ws->prefix[s]);
  const double nan = __longlong_as_double(0x7FF8000000000000ll);
  double dT, dT99, dTmax, dH2, dH2max, dCO;
  // an empty mask leaves the single value 0.0 (solver_3d.py:387-400); a NaN
  // among the values makes every figure of that array NaN, as np.percentile
  if (ws->count[0] == 0) {
    dT = dT99 = dTmax = 0.0;
  } else if (ws->nan_count[0]) {
    dT = dT99 = dTmax = nan;
  } else {
    dT = np_lerp(val[0], val[1], ws->gamma[0][0]);
    dT99 = np_lerp(val[2], val[3], ws->gamma[1][0]);
    dTmax = val[4];
  }
  if (ws->count[1] == 0) {
    dH2 = dH2max = 0.0;
  } else if (ws->nan_count[1]) {
    dH2 = dH2max = nan;
  } else {
    dH2 = np_lerp(val[5], val[6], ws->gamma[2][0]);
    dH2max = val[7];
  }
  if (ws->count[2] == 0) dCO = 0.0;
  else if (ws->nan_count[2]) dCO = nan;
  else dCO = np_lerp(val[8], val[9], ws->gamma[2][1]);
  // Python's max(dT, dH2, dCO): keeps the earlier value unless a later one is larger
  double dmax = dT;
  if (dH2 > dmax) dmax = dH2;
  if (dCO > dmax) dmax = dCO;
  const double hot = ws->count[0] == 0 ? (0.0 > tol ? 1.0 : 0.0) : (double)ws->n_hot;
  rec[0] = dT;
  rec[1] = dT99;
  rec[2] = dH2;
  rec[3] = dCO;
  rec[4] = dTmax;
  rec[5] = dH2max;
  rec[6] = hot;
  rec[7] = dmax;
  rec[8] = (double)ws->count[0];
  rec[9] = (double)ws->count[1];
  rec[10] = (double)ws->count[2];
  rec[11] = dmax < tol ? 1.0 : 0.0;
}

// ---------------------------------------------------------------------------
// Snapshot and damping
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256)
snapshot_kernel(int64_t n, int n_species, const double *__restrict__ t_gas,
                const double *__restrict__ y, double *__restrict__ t_old,
                double *__restrict__ y_old) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  t_old[i] = t_gas[i];
  for (int s = 0; s < n_species; ++s) y_old[(int64_t)s * n + i] = y[(int64_t)s * n + i];
}

__device__ __forceinline__ double floor30(double v) { return v < 1e-30 ? 1e-30 : v; }

__global__ void __launch_bounds__(256)
damp_abundances_kernel(int64_t n, int mode, int n_species, double *__restrict__ y,
                       const double *__restrict__ y_old, const double *__restrict__ w_damp,
                       int r_h, int r_h2, int r_cp, int r_c, int r_co, double x_c_total) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mode == P3D_DAMP_UNIFORM) {
    for (int s = 0; s < n_species; ++s) {
      const int64_t k = (int64_t)s * n + i;
      y[k] = 0.5 * y_old[k] + 0.5 * y[k];
    }
    return;
  }
  double *h = y + (int64_t)r_h * n, *h2 = y + (int64_t)r_h2 * n;
  double *cp = y + (int64_t)r_cp * n, *co = y + (int64_t)r_co * n;
  const double *c = y + (int64_t)r_c * n;
  const double h2o = y_old[(int64_t)r_h2 * n + i], coo = y_old[(int64_t)r_co * n + i];
  double xh2, xco;
  if (mode == P3D_DAMP_REFERENCE) {  // solver_3d.py:345-354
    xh2 = 0.5 * h2o + 0.5 * h2[i];
    xco = 0.5 * coo + 0.5 * co[i];
  } else {  // P3D_DAMP_OSCILLATING, solver_3d.py:330-345
    const double w = w_damp[i];
    if (!(w > 0.55)) return;
    xh2 = w * h2o + (1.0 - w) * h2[i];
    xco = w * coo + (1.0 - w) * co[i];
  }
  h2[i] = xh2;
  h[i] = floor30(1.0 - 2.0 * xh2);
  co[i] = xco;
  cp[i] = floor30(x_c_total - xco - c[i]);
}

__global__ void __launch_bounds__(256)
damp_temperature_kernel(int64_t n, const double *__restrict__ t_old, double *__restrict__ t_gas,
                        double *__restrict__ w_damp) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double to = t_old[i], tn = t_gas[i];
  const double rel = fabs(tn - to) / (to < 10.0 ? 10.0 : to);  // np.maximum(old_T, 10)
  double w = w_damp[i];
  w = rel > 0.5 ? w + 0.15 : (rel < 0.1 ? w - 0.02 : w);
  w = w < 0.5 ? 0.5 : w;  // np.clip(., 0.5, 0.98)
  w = w > 0.98 ? 0.98 : w;
  w_damp[i] = w;
  t_gas[i] = w * to + (1.0 - w) * tn;
}

}  // namespace

// Workspace of p3d_convergence_noalloc: one copy per GPU, held until the
// process ends.  Host code never passes this name to a kernel (on the host it
// is only a stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[P3D_CONV_WORKSPACE_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (looked up once per device).
static int scratch_address(void **out) {
  int dev = 0;
  P3D_CHECK_HIP(hipGetDevice(&dev));
  P3D_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the supported %d", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress of the convergence scratch failed: %s",
                hipGetErrorString(e));
      return P3D_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return P3D_OK;
}

}  // namespace p3d

using namespace p3d;

extern "C" {

int p3d_loop_snapshot(int64_t n, int32_t n_species, const double *t_gas, const double *y,
                      double *t_old, double *y_old, void *stream) {
  P3D_REQUIRE(n >= 0 && n_species >= 0 && n_species <= 128,
              "p3d_loop_snapshot: n >= 0 and 0 <= n_species <= 128 required");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(t_gas && t_old && (n_species == 0 || (y && y_old)),
              "p3d_loop_snapshot: null device array");
  hipLaunchKernelGGL(snapshot_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     n, (int)n_species, t_gas, y, t_old, y_old);
  return launch_status("snapshot_kernel");
}

int p3d_damp_abundances(int64_t n, int32_t mode, int32_t n_species, double *y,
                        const double *y_old, const double *w_damp, const int32_t *rows5,
                        double x_c_total, void *stream) {
  P3D_REQUIRE(n >= 0 && n_species > 0 && n_species <= 128,
              "p3d_damp_abundances: n >= 0 and 1 <= n_species <= 128 required");
  P3D_REQUIRE(mode == P3D_DAMP_REFERENCE || mode == P3D_DAMP_UNIFORM ||
                  mode == P3D_DAMP_OSCILLATING,
              "p3d_damp_abundances: unknown mode %d", (int)mode);
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(y && y_old, "p3d_damp_abundances: null device array");
  int r[5] = {0, 0, 0, 0, 0};
  if (mode != P3D_DAMP_UNIFORM) {
    P3D_REQUIRE(rows5, "p3d_damp_abundances: rows5 (H, H2, C+, C, CO) required");
    for (int k = 0; k < 5; ++k) {
      P3D_REQUIRE(rows5[k] >= 0 && rows5[k] < n_species,
                  "p3d_damp_abundances: row %d out of range", (int)rows5[k]);
      r[k] = rows5[k];
    }
    P3D_REQUIRE(mode != P3D_DAMP_OSCILLATING || w_damp,
                "p3d_damp_abundances: w_damp required for the oscillating-cell mode");
  }
  hipLaunchKernelGGL(damp_abundances_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     as_stream(stream), n, (int)mode, (int)n_species, y, y_old, w_damp, r[0],
                     r[1], r[2], r[3], r[4], x_c_total);
  return launch_status("damp_abundances_kernel");
}

int p3d_damp_temperature(int64_t n, const double *t_old, double *t_gas, double *w_damp,
                         void *stream) {
  P3D_REQUIRE(n >= 0, "p3d_damp_temperature: n >= 0 required");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(t_old && t_gas && w_damp, "p3d_damp_temperature: null device array");
  hipLaunchKernelGGL(damp_temperature_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     as_stream(stream), n, t_old, t_gas, w_damp);
  return launch_status("damp_temperature_kernel");
}

int p3d_convergence(int64_t n, const double *t_old, const double *t_gas, const double *h2_old,
                    const double *x_h2, const double *co_old, const double *x_co, double tol,
                    double *record, void *workspace, void *stream) {
  P3D_REQUIRE(n >= 0 && n <= ((int64_t)1 << 40), "p3d_convergence: 0 <= n <= 2^40 required");
  P3D_REQUIRE(record && workspace, "p3d_convergence: record and workspace required");
  P3D_REQUIRE(n == 0 || (t_old && t_gas && h2_old && x_h2 && co_old && x_co),
              "p3d_convergence: null device array");
  P3D_REQUIRE(((uintptr_t)workspace & 7u) == 0, "p3d_convergence: workspace must be 8-byte aligned");
  hipStream_t s = as_stream(stream);
  ConvWs *ws = static_cast<ConvWs *>(workspace);
  hipLaunchKernelGGL(conv_reset, dim3(1), dim3(64), 0, s, ws);
  const unsigned grid = (unsigned)std::min<int64_t>(kHistGrid, std::max<int64_t>(1, (n + kHistBlock - 1) / kHistBlock));
  for (int p = 0; p < kPasses; ++p) {
    hipLaunchKernelGGL(conv_hist, dim3(grid), dim3(kHistBlock), 0, s, n, p, tol, t_old, t_gas,
                       h2_old, x_h2, co_old, x_co, ws);
    hipLaunchKernelGGL(conv_step, dim3(1), dim3(64 * kSel), 0, s, p, ws);
  }
  hipLaunchKernelGGL(conv_final, dim3(1), dim3(64), 0, s, tol, ws, record);
  return launch_status("convergence kernels");
}

int p3d_convergence_noalloc(int64_t n, const double *t_old, const double *t_gas,
                            const double *h2_old, const double *x_h2, const double *co_old,
                            const double *x_co, double tol, double *record, void *stream) {
  void *ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != P3D_OK) return rc;
  return p3d_convergence(n, t_old, t_gas, h2_old, x_h2, co_old, x_co, tol, record, ws, stream);
}

}  // extern "C"
