// Shared helpers of the gfx950 kernels: error reporting, stream casts,
// physical constants and 64-lane wavefront collectives.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/prism3d_amd.h"

namespace p3d {

// --- error reporting (per calling thread) ---------------------------------
void set_error(const char *fmt, ...);

#define P3D_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::p3d::set_error("%s failed: %s", #expr, hipGetErrorString(e_));       \
      return P3D_EHIP;                                                        \
    }                                                                         \
  } while (0)

#define P3D_REQUIRE(cond, ...)                                                \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::p3d::set_error(__VA_ARGS__);                                          \
      return P3D_EINVAL;                                                      \
    }                                                                         \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Launch-error check after a kernel launch.
inline int launch_status(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return P3D_EHIP;
  }
  return P3D_OK;
}

inline unsigned grid_for(int64_t n, int block) {
  return static_cast<unsigned>((n + block - 1) / block);
}

// --- constants (CGS; upstream prism3d/utils/constants.py) ------------------
constexpr double kC = 2.99792458e10;
constexpr double kH = 6.62607015e-27;
constexpr double kK = 1.380649e-16;
constexpr double kMH = 1.6735575e-24;
constexpr double kEV = 1.602176634e-12;
constexpr double kPi = 3.14159265358979323846;

// --- wavefront collectives (64 lanes) ---------------------------------------
// Butterfly reductions: every lane ends with the same bits (each combine is
// commutative), and readfirstlane makes the value wave-uniform for the
// compiler so that branches on it are scalar branches.
__device__ __forceinline__ double uniform(double v) {
  int2 b = *reinterpret_cast<int2 *>(&v);
  b.x = __builtin_amdgcn_readfirstlane(b.x);
  b.y = __builtin_amdgcn_readfirstlane(b.y);
  return *reinterpret_cast<double *>(&b);
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return uniform(v);
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return uniform(v);
}
// Value held by lane `src` (src wave-uniform), broadcast to all lanes.
__device__ __forceinline__ double bcast(double v, int src) {
  int2 b = *reinterpret_cast<int2 *>(&v);
  b.x = __builtin_amdgcn_readlane(b.x, src);
  b.y = __builtin_amdgcn_readlane(b.y, src);
  return *reinterpret_cast<double *>(&b);
}
// Largest |v| over lanes and the lowest lane holding it.
__device__ __forceinline__ int wave_argmax(double v, int lane) {
  int idx = lane;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    double ov = __shfl_xor(v, m, 64);
    int oi = __shfl_xor(idx, m, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  return uniform(idx);
}

}  // namespace p3d
