// Tree-ensemble chemistry surrogate: the inference of the upstream ML
// accelerator (prism3d/chemistry/accelerator.py:346-411, used in place of the
// explicit chemistry by accelerator.py:472-510).  Per cell: seven log
// features with floors, standardised and rounded to float32 as the tree
// library does; for each of the nine species one tree ensemble (gradient
// boosting: base + sum of scale * leaf, stage by stage; random forest: sum of
// leaves / number of trees), de-standardised and raised to 10^y; then H and
// C renormalised, O by difference and e- from the cations.
//
// Trees are stored as perfect binary trees of one depth D (breadth-first,
// children of node i at 2i+1 and 2i+2; shallower leaves replicated), so a
// walk is D dependent loads with no branch: the wavefront stays converged.
// An internal node is one 64-bit word {float threshold, int feature}: the
// library compares float32 features with float64 thresholds, which equals a
// float32 comparison against the threshold rounded toward -inf (done on the
// host).  Four trees are walked together so four load chains are in flight.
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "p3d_common.h"

namespace p3d {

constexpr int kSurOut = 9;   // H, H2, C+, C, CO, O, e-, OH, HCO+
constexpr int kSurFeat = 7;  // log nH, log T, log G0, A_V, log zeta, log fsh_H2, log fsh_CO
constexpr int kSurMaxDepth = 16;

struct SurrogateDev {
  int n_trees[kSurOut], tree0[kSurOut];
  double base[kSurOut], scale[kSurOut], divisor[kSurOut];
  double x_mean[kSurFeat], x_scale[kSurFeat], y_mean[kSurOut], y_scale[kSurOut];
  const uint2 *node;   // [tree][2^D - 1]
  const double *leaf;  // [tree][2^D]
};

struct SurrogateOut {
  double *p[kSurOut];
};

struct Surrogate {
  SurrogateDev d;
  int depth = 0;
  void *blob = nullptr;
  ~Surrogate() {
    if (blob) (void)hipFree(blob);
  }
};

static std::mutex g_sur_mu;
static std::map<int64_t, std::unique_ptr<Surrogate>> g_surs;
static int64_t g_next_sur = 1;

__device__ __forceinline__ float pick_feature(const float (&x)[kSurFeat], int f) {
  float v = x[0];
#pragma unroll
  for (int k = 1; k < kSurFeat; ++k) v = f == k ? x[k] : v;
  return v;
}

// leaf index of the walk of one perfect tree of depth D
template <int D>
__device__ __forceinline__ int walk(const uint2 *__restrict__ nd, const float (&x)[kSurFeat]) {
  int i = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const uint2 w = nd[i];
    const float v = pick_feature(x, (int)w.y);
    i = 2 * i + (v <= __uint_as_float(w.x) ? 1 : 2);  // NaN goes right, as the library
  }
  return i - ((1 << D) - 1);
}

template <int D>
__global__ void __launch_bounds__(256)
surrogate_kernel(int64_t n, const double *__restrict__ nh, const double *__restrict__ tg,
                 const double *__restrict__ g0, const double *__restrict__ av,
                 const double *__restrict__ zeta, const double *__restrict__ fh2,
                 const double *__restrict__ fco, const SurrogateDev S, double xc_tot,
                 double xo_tot, const SurrogateOut out) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int NI = (1 << D) - 1, NL = 1 << D;
  const double f[kSurFeat] = {log10(fmax(nh[i], 1.0)),    log10(fmax(tg[i], 5.0)),
                              log10(fmax(g0[i], 1e-2)),   av[i],
                              log10(fmax(zeta[i], 1e-18)), log10(fmax(fh2[i], 1e-10)),
                              log10(fmax(fco[i], 1e-10))};
  float x[kSurFeat];
#pragma unroll
  for (int k = 0; k < kSurFeat; ++k) x[k] = (float)((f[k] - S.x_mean[k]) / S.x_scale[k]);
  double r[kSurOut];
#pragma unroll
  for (int o = 0; o < kSurOut; ++o) {
    double raw = S.base[o];
    const double sc = S.scale[o];
    const int nt = S.n_trees[o];
    const uint2 *nd = S.node + (int64_t)S.tree0[o] * NI;
    const double *lf = S.leaf + (int64_t)S.tree0[o] * NL;
    int t = 0;
    for (; t + 4 <= nt; t += 4) {
      const int a = walk<D>(nd + (int64_t)t * NI, x);
      const int b = walk<D>(nd + (int64_t)(t + 1) * NI, x);
      const int c = walk<D>(nd + (int64_t)(t + 2) * NI, x);
      const int e = walk<D>(nd + (int64_t)(t + 3) * NI, x);
      const double va = lf[(int64_t)t * NL + a], vb = lf[(int64_t)(t + 1) * NL + b];
      const double vc = lf[(int64_t)(t + 2) * NL + c], ve = lf[(int64_t)(t + 3) * NL + e];
      raw += sc * va;  // stage order of the library's accumulation
      raw += sc * vb;
      raw += sc * vc;
      raw += sc * ve;
    }
    for (; t < nt; ++t) raw += sc * lf[(int64_t)t * NL + walk<D>(nd + (int64_t)t * NI, x)];
    raw = raw / S.divisor[o];
    r[o] = pow(10.0, raw * S.y_scale[o] + S.y_mean[o]);
  }
  // conservation of the batch prediction (accelerator.py:388-409)
  const double hsum = r[0] + 2.0 * r[1];
  const double hs = hsum > 0.0 ? 1.0 / hsum : 1.0;
  r[0] = r[0] * hs;
  r[1] = r[1] * hs;
  const double csum = r[2] + r[3] + r[4];
  const double cs = csum > 0.0 ? xc_tot / csum : 1.0;
  r[2] = r[2] * cs;
  r[3] = r[3] * cs;
  r[4] = r[4] * cs;
  r[5] = fmax(xo_tot - r[4] - r[7], 1e-30);
  r[6] = r[2] + r[8];
#pragma unroll
  for (int o = 0; o < kSurOut; ++o) out.p[o][i] = r[o];
}

template <int D>
static void launch_depth(hipStream_t s, int64_t n, const double *const *in, const SurrogateDev &S,
                         double xc, double xo, const SurrogateOut &o) {
  hipLaunchKernelGGL(surrogate_kernel<D>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, in[0],
                     in[1], in[2], in[3], in[4], in[5], in[6], S, xc, xo, o);
}

static int launch_surrogate(int depth, hipStream_t s, int64_t n, const double *const *in,
                            const SurrogateDev &S, double xc, double xo, const SurrogateOut &o) {
  switch (depth) {
#define P3D_SUR_CASE(D) \
  case D:               \
    launch_depth<D>(s, n, in, S, xc, xo, o); break;
    P3D_SUR_CASE(1) P3D_SUR_CASE(2) P3D_SUR_CASE(3) P3D_SUR_CASE(4) P3D_SUR_CASE(5)
    P3D_SUR_CASE(6) P3D_SUR_CASE(7) P3D_SUR_CASE(8) P3D_SUR_CASE(9) P3D_SUR_CASE(10)
    P3D_SUR_CASE(11) P3D_SUR_CASE(12) P3D_SUR_CASE(13) P3D_SUR_CASE(14) P3D_SUR_CASE(15)
    P3D_SUR_CASE(16)
#undef P3D_SUR_CASE
    default:
      set_error("tree depth %d outside [1, %d]", depth, kSurMaxDepth);
      return P3D_EINVAL;
  }
  return launch_status("surrogate_kernel");
}

}  // namespace p3d

using namespace p3d;

extern "C" {

int p3d_surrogate_create(int32_t n_outputs, int32_t n_features, int32_t depth,
                         const int32_t *n_trees, const double *base, const double *scale,
                         const double *divisor, const int32_t *feature, const float *threshold,
                         const double *leaf, const double *x_mean, const double *x_scale,
                         const double *y_mean, const double *y_scale, int64_t *handle) {
  P3D_REQUIRE(handle, "handle pointer is required");
  P3D_REQUIRE(n_outputs == kSurOut && n_features == kSurFeat,
              "the surrogate maps %d features to %d species (got %d -> %d)", kSurFeat, kSurOut,
              n_features, n_outputs);
  P3D_REQUIRE(depth >= 1 && depth <= kSurMaxDepth, "depth must be in [1, %d] (got %d)",
              kSurMaxDepth, depth);
  P3D_REQUIRE(n_trees && base && scale && divisor && feature && threshold && leaf && x_mean &&
                  x_scale && y_mean && y_scale,
              "null host array");
  int64_t total = 0;
  for (int o = 0; o < kSurOut; ++o) {
    P3D_REQUIRE(n_trees[o] >= 0 && n_trees[o] <= (1 << 20), "n_trees[%d] out of range", o);
    P3D_REQUIRE(divisor[o] != 0.0, "divisor[%d] must be non-zero", o);
    total += n_trees[o];
  }
  const int64_t ni = (1LL << depth) - 1, nl = 1LL << depth;
  P3D_REQUIRE(total * nl * 16 <= (int64_t)1 << 34, "tree ensemble larger than 16 GiB");
  for (int64_t k = 0; k < total * ni; ++k)
    P3D_REQUIRE(feature[k] >= 0 && feature[k] < kSurFeat, "feature index %d at node %lld",
                feature[k], (long long)k);
  auto sur = std::make_unique<Surrogate>();
  sur->depth = depth;
  SurrogateDev &d = sur->d;
  int64_t t0 = 0;
  for (int o = 0; o < kSurOut; ++o) {
    d.n_trees[o] = n_trees[o];
    d.tree0[o] = (int)t0;
    t0 += n_trees[o];
    d.base[o] = base[o];
    d.scale[o] = scale[o];
    d.divisor[o] = divisor[o];
    d.y_mean[o] = y_mean[o];
    d.y_scale[o] = y_scale[o];
  }
  for (int k = 0; k < kSurFeat; ++k) {
    d.x_mean[k] = x_mean[k];
    d.x_scale[k] = x_scale[k];
  }
  const size_t node_bytes = (size_t)(total * ni) * sizeof(uint2);
  const size_t leaf_off = (node_bytes + 255) & ~(size_t)255;
  const size_t bytes = leaf_off + (size_t)(total * nl) * sizeof(double);
  std::vector<char> host(bytes, 0);
  uint2 *hn = reinterpret_cast<uint2 *>(host.data());
  for (int64_t k = 0; k < total * ni; ++k) {
    uint32_t bits;
    std::memcpy(&bits, &threshold[k], 4);
    hn[k] = make_uint2(bits, (uint32_t)feature[k]);
  }
  std::memcpy(host.data() + leaf_off, leaf, (size_t)(total * nl) * sizeof(double));
  if (bytes > 0) {
    if (hipMalloc(&sur->blob, bytes) != hipSuccess) {
      set_error("hipMalloc of %zu bytes for the tree ensemble failed", bytes);
      return P3D_ENOMEM;
    }
    P3D_CHECK_HIP(hipMemcpy(sur->blob, host.data(), bytes, hipMemcpyHostToDevice));
  }
  d.node = reinterpret_cast<const uint2 *>(sur->blob);
  d.leaf = reinterpret_cast<const double *>(static_cast<char *>(sur->blob) + leaf_off);
  std::lock_guard<std::mutex> lk(g_sur_mu);
  const int64_t h = g_next_sur++;
  g_surs[h] = std::move(sur);
  *handle = h;
  return P3D_OK;
}

int p3d_surrogate_destroy(int64_t handle) {
  std::lock_guard<std::mutex> lk(g_sur_mu);
  auto it = g_surs.find(handle);
  P3D_REQUIRE(it != g_surs.end(), "unknown surrogate handle %lld", (long long)handle);
  g_surs.erase(it);
  return P3D_OK;
}

int p3d_surrogate_predict(int64_t handle, int64_t n, const double *n_h, const double *t_gas,
                          const double *g0, const double *a_v, const double *zeta,
                          const double *fsh_h2, const double *fsh_co, double x_c_total,
                          double x_o_total, double *const *out, void *stream) {
  Surrogate *sur = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_sur_mu);
    auto it = g_surs.find(handle);
    if (it != g_surs.end()) sur = it->second.get();
  }
  P3D_REQUIRE(sur, "unknown surrogate handle %lld", (long long)handle);
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(out, "output pointer array (host, %d device pointers) is required", kSurOut);
  P3D_REQUIRE(n_h && t_gas && g0 && a_v && zeta && fsh_h2 && fsh_co, "null device array");
  SurrogateOut o;
  for (int k = 0; k < kSurOut; ++k) {
    P3D_REQUIRE(out[k], "null output array %d", k);
    o.p[k] = out[k];
  }
  const double *in[kSurFeat] = {n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co};
  return launch_surrogate(sur->depth, as_stream(stream), n, in, sur->d, x_c_total, x_o_total, o);
}

}  // extern "C"
