// Persistent one-cell-per-lane chemistry solve on gfx950.
//
// Each lane of every resident wavefront runs the trip-stepped per-cell solve
// of lane_bdf.h on its own cell and, when the cell is done, takes the next one
// from a device work counter (one 64-bit atomic per cell), so cells whose
// step counts differ by orders of magnitude do not leave lanes idle.  A
// wavefront exits when all 64 lanes have found the queue empty.  Per-cell
// arrays live in slot memory (entry-major, coalesced across the wavefront);
// the network's RHS, Jacobian and statically ordered sparse LU are the
// straight-line code of lane_net.h.
#define P3D_HD __host__ __device__
#define P3D_INLINE __forceinline__
#define P3D_ANY(x) (__any((x) ? 1 : 0) != 0)
#include "p3d_common.h"
#include "lane_bdf.h"
#include "lane_launch.h"

namespace p3d {

namespace {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Slot memory through one buffer resource per wavefront.  Wavefront w owns
// one contiguous block of kSlotSize x 64 doubles, entry-major: entry e of
// lane l at byte e * 512 + l * 8 of the block.  A wavefront's access to one
// entry is one contiguous 512-byte segment (four cache lines), its working
// set spans a few pages, and entry offsets are compile-time constants.  (A
// lane-major block, each lane's entries contiguous, measured faster only
// while most blocks ran with few active lanes.)  The resource's record count
// is the wavefront's block, so the hardware range check confines every access
// to it.
constexpr unsigned kEntryStride = 64 * sizeof(double);
constexpr unsigned kWaveBlock = lane::kSlotSize * kEntryStride;

struct SlotBuf {
  __amdgpu_buffer_rsrc_t r;
  unsigned voff;
  struct Ref {
    const SlotBuf *m;
    int e;
    __device__ __forceinline__ operator double() const {
      return __builtin_bit_cast(
          double, __builtin_amdgcn_raw_buffer_load_b64(m->r, m->voff, (unsigned)e * kEntryStride, 0));
    }
    __device__ __forceinline__ const Ref &operator=(double v) const {
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), m->r, m->voff,
                                            (unsigned)e * kEntryStride, 0);
      return *this;
    }
    // entry-to-entry assignment stores the value (the implicit copy
    // assignment would copy the proxy instead)
    __device__ __forceinline__ const Ref &operator=(const Ref &o) const {
      return *this = (double)o;
    }
  };
  __device__ __forceinline__ Ref at(int e) const { return Ref{this, e}; }
  __device__ __forceinline__ Ref k(int q) const { return at(lane::kOffK + q); }
  __device__ __forceinline__ Ref J(int z) const { return at(lane::kOffJ + z); }
  __device__ __forceinline__ Ref LU(int z) const { return at(lane::kOffLU + z); }
  __device__ __forceinline__ Ref D(int j, int i) const { return at(lane::kOffD + j * lane::kNS + i); }
  __device__ __forceinline__ Ref yp(int i) const { return at(lane::kOffYP + i); }
  __device__ __forceinline__ Ref psi(int i) const { return at(lane::kOffPSI + i); }
  __device__ __forceinline__ Ref yc(int i) const { return at(lane::kOffYC + i); }
  __device__ __forceinline__ Ref yn(int i) const { return at(lane::kOffYN + i); }
  __device__ __forceinline__ Ref dd(int i) const { return at(lane::kOffDD + i); }
};

struct DevSrc {
  int64_t n;
  unsigned long long *ctr;
  unsigned *exited, *n_rescue;
  int64_t *rescue;
  unsigned exit_threshold;
  double *y;
  const double *nH, *T, *G0, *AV, *zeta, *fH2, *fCO;
  int32_t *conv, *stats;

  __device__ int64_t fetch() {
    const unsigned long long c = atomicAdd(ctr, 1ull);
    if (c < (unsigned long long)n) return (int64_t)c;
    atomicAdd(exited, 1u);
    return -1;
  }
  // true once enough lanes found the queue empty: the remaining cells are
  // the long tail and go to the one-cell-per-workgroup kernel
  __device__ bool should_abandon() const {
    return __hip_atomic_load(exited, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
           exit_threshold;
  }
  __device__ void abandon(int64_t c) {
    const unsigned k = atomicAdd(n_rescue, 1u);
    rescue[k] = c;
  }
  __device__ void load(int64_t c, lane::Cell &p, double (&yv)[lane::kNS]) {
    p.nH = nH[c]; p.T = T[c]; p.G0 = G0[c]; p.AV = AV[c];
    p.zeta = zeta[c]; p.fH2 = fH2[c]; p.fCO = fCO[c];
#pragma unroll
    for (int s = 0; s < lane::kNS; ++s) yv[s] = y[(int64_t)s * n + c];
  }
  __device__ void store(int64_t c, const double (&yv)[lane::kNS], int cv, const int (&st)[4]) {
#pragma unroll
    for (int s = 0; s < lane::kNS; ++s) y[(int64_t)s * n + c] = yv[s];
    if (conv) conv[c] = cv;
    if (stats)
#pragma unroll
      for (int k = 0; k < 4; ++k) stats[4 * c + k] = st[k];
  }
};

// Makes every 32-bit word of a (wave-uniform) argument struct opaque to the
// optimiser for one trip, so nothing derived from it is loop invariant.
template <class T>
__device__ __forceinline__ void launder(T &x) {
  static_assert(sizeof(T) % 4 == 0, "launder needs whole dwords");
  unsigned *w = reinterpret_cast<unsigned *>(&x);
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) __asm__ __volatile__("" : "+s"(w[i]));
}

}  // namespace

__global__ void __launch_bounds__(64)
lane_counter_reset(unsigned long long *head, unsigned *exited, unsigned *n_rescue) {
  if (threadIdx.x == 0) {
    *head = 0ull;
    *exited = 0u;
    *n_rescue = 0u;
  }
}

// W: requested wavefronts per SIMD (register budget 512 / W); the spill the
// compiler then needs is small, and the choice is a tuning parameter.
template <int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W, W)))
chem_lane_kernel(lane::Tables tb, lane::Opts o, DevSrc src, double *__restrict__ ws) {
  SlotBuf m;
  m.r = __builtin_amdgcn_make_buffer_rsrc(ws + (size_t)blockIdx.x * lane::kSlotSize * 64, (short)0,
                                          (int)kWaveBlock, 0x00020000);
  m.voff = threadIdx.x * 8u;
  lane::Lane L;
  lane::lane_init(L);
  for (;;) {
    // Everything derived from the arguments (slot entry offsets, the input
    // array's species offsets, tolerance products, log10 t_start ...) is loop
    // invariant; the compiler would hoist hundreds of such values out of the
    // loop and keep them live in registers.  Laundering the arguments each
    // trip makes it recompute them where they are used.
    SlotBuf mt = m;
    lane::Tables tt = tb;
    lane::Opts ot = o;
    DevSrc st = src;
    launder(tt);
    launder(ot);
    launder(st);
    lane::trip(L, mt, tt, ot, st);
    if (__all(L.phase == lane::kExit)) break;
  }
}

bool lane_network_matches(int ns, int nr, const int *rtype, const int *reactants,
                          const int *products, int e_index, const int *roles7) {
  if (ns != lane::kNS || nr != lane::kNR || e_index != lane::kE) return false;
  for (int r = 0; r < nr; ++r) {
    if (rtype[r] != lane::kRtype[r]) return false;
    for (int k = 0; k < 2; ++k)
      if (reactants[2 * r + k] != lane::kReact[2 * r + k]) return false;
    for (int k = 0; k < 3; ++k)
      if (products[3 * r + k] != lane::kProd[3 * r + k]) return false;
  }
  for (int k = 0; k < 7; ++k)
    if (roles7[k] != lane::kRoles[k]) return false;
  return true;
}

int lane_slot_doubles() { return lane::kSlotSize; }

int lane_launch(const LaneArgs &a, int64_t n_waves, int occupancy, double *ws,
                const LaneQueue &wq) {
  const p3d_solve_opts &q = a.opts;
  lane::Opts o;
  o.t_max = q.t_max; o.t_start = q.t_start; o.rtol = q.rtol; o.atol = q.atol;
  o.stop_rtol = q.stop_rtol; o.x_c_total = q.x_c_total; o.x_o_total = q.x_o_total;
  o.newton_tol = fmax(10.0 * 2.220446049250313e-16 / q.rtol, fmin(0.03, sqrt(q.rtol)));
  o.nk = q.n_checkpoints; o.early_stop = q.early_stop; o.post_mode = q.post_mode;
  o.max_steps = q.max_steps;
  o.rate_fp32 = q.rate_fp32;
  lane::Tables tb{a.rtype, a.alpha, a.beta, a.gamma, a.post_w};
  DevSrc src{a.n,     wq.head, wq.exited, wq.n_rescue, wq.rescue, wq.exit_threshold,
             a.y,     a.nH,    a.T,      a.G0,       a.AV,     a.zeta,
             a.fH2,   a.fCO,   a.conv,   a.stats};
  hipStream_t s = as_stream(a.stream);
  hipLaunchKernelGGL(lane_counter_reset, dim3(1), dim3(64), 0, s, wq.head, wq.exited,
                     wq.n_rescue);
  const dim3 grid((unsigned)n_waves), block(64);
  if (occupancy == 4)
    hipLaunchKernelGGL(chem_lane_kernel<4>, grid, block, 0, s, tb, o, src, ws);
  else if (occupancy == 3)
    hipLaunchKernelGGL(chem_lane_kernel<3>, grid, block, 0, s, tb, o, src, ws);
  else
    hipLaunchKernelGGL(chem_lane_kernel<2>, grid, block, 0, s, tb, o, src, ws);
  return launch_status("chem_lane_kernel");
}

}  // namespace p3d
