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
// Reductions run in DPP within each 16-lane row (quad swaps, half-row and
// row mirrors), then combine the four row results with readlane into a
// wave-uniform scalar.  The combine order is fixed, so every call returns the
// same bits on every lane; the scalar result also makes branches on it
// scalar branches.  All 64 lanes must be active.
__device__ __forceinline__ double uniform(double v) {
  int2 b = *reinterpret_cast<int2 *>(&v);
  b.x = __builtin_amdgcn_readfirstlane(b.x);
  b.y = __builtin_amdgcn_readfirstlane(b.y);
  return *reinterpret_cast<double *>(&b);
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Value held by lane `src` (src wave-uniform), broadcast to all lanes.
__device__ __forceinline__ double bcast(double v, int src) {
  int2 b = *reinterpret_cast<int2 *>(&v);
  b.x = __builtin_amdgcn_readlane(b.x, src);
  b.y = __builtin_amdgcn_readlane(b.y, src);
  return *reinterpret_cast<double *>(&b);
}

template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  int2 b = *reinterpret_cast<int2 *>(&v);
  b.x = __builtin_amdgcn_update_dpp(0, b.x, CTRL, 0xF, 0xF, false);
  b.y = __builtin_amdgcn_update_dpp(0, b.y, CTRL, 0xF, 0xF, false);
  return *reinterpret_cast<double *>(&b);
}
template <int CTRL>
__device__ __forceinline__ int dpp(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
// DPP controls: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp<kDppXor1>(v);
  v += dpp<kDppXor2>(v);
  v += dpp<kDppHalfMirror>(v);
  v += dpp<kDppMirror>(v);
  return (bcast(v, 0) + bcast(v, 16)) + (bcast(v, 32) + bcast(v, 48));
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp<kDppXor1>(v));
  v = fmax(v, dpp<kDppXor2>(v));
  v = fmax(v, dpp<kDppHalfMirror>(v));
  v = fmax(v, dpp<kDppMirror>(v));
  return fmax(fmax(bcast(v, 0), bcast(v, 16)), fmax(bcast(v, 32), bcast(v, 48)));
}

template <int CTRL>
__device__ __forceinline__ void argmax_step(double &v, int &idx) {
  const double ov = dpp<CTRL>(v);
  const int oi = dpp<CTRL>(idx);
  if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
}
// Largest v over lanes and the lowest lane holding it (wave-uniform).
__device__ __forceinline__ int wave_argmax(double v, int lane) {
  int idx = lane;
  argmax_step<kDppXor1>(v, idx);
  argmax_step<kDppXor2>(v, idx);
  argmax_step<kDppHalfMirror>(v, idx);
  argmax_step<kDppMirror>(v, idx);
  double bv = bcast(v, 0);
  int bi = __builtin_amdgcn_readlane(idx, 0);
#pragma unroll
  for (int r = 16; r < 64; r += 16) {
    const double rv = bcast(v, r);
    const int ri = __builtin_amdgcn_readlane(idx, r);
    if (rv > bv || (rv == bv && ri < bi)) { bv = rv; bi = ri; }
  }
  return uniform(bi);
}

}  // namespace p3d
