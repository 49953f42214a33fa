// Tree-ensemble chemistry surrogate: the inference of the upstream ML
// accelerator (prism3d/chemistry/accelerator.py:346-411, used in place of the
// explicit chemistry by accelerator.py:472-510).  Per cell: seven log
// features with floors, standardised and rounded to float32 as the tree
// library does; for each of the nine species one tree ensemble (gradient
// boosting: base + sum of scale * leaf, stage by stage; random forest: sum of
// leaves / number of trees), de-standardised and raised to 10^y; then H and
// C renormalised, O by difference and e- from the cations.
//
// Trees arrive as perfect binary trees of one depth D (breadth-first,
// children of node i at 2i+1 and 2i+2; shallower leaves replicated), so a
// walk has no branch and the wavefront stays converged.  The library
// compares float32 features with float64 thresholds, which equals a float32
// comparison against the threshold rounded toward -inf (done by the caller).
//
// Rank encoding.  For each feature the library builds the sorted list of the
// distinct thresholds the ensemble tests it against; a cell's feature value
// x becomes its rank r = #{t < x} (a binary search, once per cell), and a
// node testing x <= t_j becomes r <= j.  The two are equivalent for every
// non-NaN x; NaN gets a rank above every j and goes right, as the library.
// A node is then one 32-bit word, feature in the top 3 bits and j below.
//
// Two-level blocks.  A node at even depth and its two children share one
// 16-byte block {node, left child, right child, 0}, so one vector load
// serves two levels: a depth-6 walk is three block loads and one leaf load
// instead of six node loads and a leaf load.  The walk is bound by the
// rate at which the texture addresser takes lane-divergent addresses, so
// fewer load instructions is the lever.  The cell's seven ranks sit in LDS
// (lane-major, no bank conflicts) and each node's lookup is one LDS read.
// Four trees are walked together so four load chains are in flight.
#include <algorithm>
#include <cmath>
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

constexpr int kSurRankBits = 29;                          // node word: j below bit 29
constexpr uint32_t kSurRankMask = (1u << kSurRankBits) - 1u;
constexpr int kSurBlock = 256;                           // threads per workgroup

struct SurrogateDev {
  int n_trees[kSurOut], tree0[kSurOut];
  double base[kSurOut], scale[kSurOut], divisor[kSurOut];
  double x_mean[kSurFeat], x_scale[kSurFeat], y_mean[kSurOut], y_scale[kSurOut];
  int thr0[kSurFeat], n_thr[kSurFeat];  // each feature's sorted distinct thresholds
  const float *thr;
  const uint4 *block;  // [tree][blocks of depth D], see block_count
  const double *leaf;  // [tree][2^D]
};

// two-level blocks of a depth-D tree: one per node at even depth
__host__ __device__ constexpr int block_count(int D) {
  int n = 0;
  for (int d = 0; d < D; d += 2) n += 1 << d;
  return n;
}

struct SurrogateOut {
  double *p[kSurOut];
};

// ---------------------------------------------------------------------------
// Multilayer-perceptron form (the accelerator's 'mlp' model type,
// accelerator.py:267-272: one regressor per species, ReLU hidden layers,
// identity output, float64 throughout).  Every layer is a small GEMM over a
// tile of 16 cells, so it runs on FP64 MFMA: a workgroup of four wavefronts
// takes 16 cells, keeps their activations in LDS (16 x width), and splits
// each layer's 16-wide output tiles over its wavefronts; one
// v_mfma_f64_16x16x4f64 per tile and 4 inputs, the bias as the initial
// accumulator, so every output sums bias, then the inputs in order.
// Widths are padded with zero weights to multiples of 16 (outputs) and 4
// (inputs); padded units stay 0 through the ReLU.
constexpr int kMlpMaxLayers = 5;    // up to 4 hidden layers + the output layer
constexpr int kMlpMaxWidth = 128;   // widest padded hidden layer
constexpr int kMlpLd = kMlpMaxWidth + 4;  // LDS row stride (doubles): 2-way bank use
constexpr int kMlpRows = 16;        // cells per workgroup
constexpr int kMlpWaves = 4;

struct MlpDev {
  int n_layers;
  int kpad[kMlpMaxLayers], npad[kMlpMaxLayers];  // padded in / out widths
  int64_t woff[kMlpMaxLayers], boff[kMlpMaxLayers];  // within one species' block
  int64_t blk;                                      // doubles per species
  const double *w;
};

struct Surrogate {
  SurrogateDev d;
  MlpDev m;
  bool mlp = false;
  int depth = 0;
  void *blob = nullptr;
  ~Surrogate() {
    if (blob) (void)hipFree(blob);
  }
};

typedef double p3d_mlp_f64x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(64 * kMlpWaves)
surrogate_mlp_kernel(int64_t n, const double *__restrict__ nh, const double *__restrict__ tg,
                     const double *__restrict__ g0, const double *__restrict__ av,
                     const double *__restrict__ zeta, const double *__restrict__ fh2,
                     const double *__restrict__ fco, const SurrogateDev S, const MlpDev M,
                     double xc_tot, double xo_tot, const SurrogateOut out) {
#pragma clang fp contract(off)
  __shared__ double feat[kMlpRows * 8];
  __shared__ double hid[2][kMlpRows * kMlpLd];
  __shared__ double yres[kSurOut][kMlpRows];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15, lq = lane >> 4;
  const int64_t c0 = (int64_t)blockIdx.x * kMlpRows;
  if (t < kMlpRows) {
    // rows past the end repeat the last cell; they are not stored
    const int64_t i = c0 + t < n ? c0 + t : n - 1;
    const double f[kSurFeat] = {log10(fmax(nh[i], 1.0)),    log10(fmax(tg[i], 5.0)),
                                log10(fmax(g0[i], 1e-2)),   av[i],
                                log10(fmax(zeta[i], 1e-18)), log10(fmax(fh2[i], 1e-10)),
                                log10(fmax(fco[i], 1e-10))};
#pragma unroll
    for (int k = 0; k < 8; ++k)
      feat[t * 8 + k] = k < kSurFeat ? (f[k] - S.x_mean[k]) / S.x_scale[k] : 0.0;
  }
  __syncthreads();
  for (int o = 0; o < kSurOut; ++o) {
    const double *wo = M.w + (int64_t)o * M.blk;
    for (int l = 0; l < M.n_layers; ++l) {
      const double *ain = l == 0 ? feat : hid[(l - 1) & 1];
      const int lda = l == 0 ? 8 : kMlpLd;
      double *aout = hid[l & 1];
      const double *W = wo + M.woff[l], *B = wo + M.boff[l];
      const int K = M.kpad[l], N = M.npad[l];
      const bool last = l == M.n_layers - 1;
      for (int nt = wave; nt < N / 16; nt += kMlpWaves) {  // wave-uniform
        const double bias = B[nt * 16 + lr];
        p3d_mlp_f64x4 acc = {bias, bias, bias, bias};
        const double *wc = W + nt * 16 + lr;
#pragma unroll 4
        for (int k0 = 0; k0 < K; k0 += 4) {
          const double a = ain[lr * lda + k0 + lq];
          const double b = wc[(int64_t)(k0 + lq) * N];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        if (!last) {
#pragma unroll
          for (int r = 0; r < 4; ++r) aout[(4 * r + lq) * kMlpLd + nt * 16 + lr] = fmax(acc[r], 0.0);
        } else if (lr == 0) {  // output column 0 of tile 0: rows 4r + lq
#pragma unroll
          for (int r = 0; r < 4; ++r) yres[o][4 * r + lq] = acc[r];
        }
      }
      __syncthreads();
    }
  }
  if (t < kMlpRows && c0 + t < n) {
    const int64_t i = c0 + t;
    double r[kSurOut];
#pragma unroll
    for (int o = 0; o < kSurOut; ++o) r[o] = pow(10.0, yres[o][t] * S.y_scale[o] + S.y_mean[o]);
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
    for (int k = 0; k < kSurOut; ++k) out.p[k][i] = r[k];
  }
}

static std::mutex g_sur_mu;
static std::map<int64_t, std::unique_ptr<Surrogate>> g_surs;
static int64_t g_next_sur = 1;

// does the node word w send this lane left?  rk: the workgroup's rank table
// (feature-major, one entry per thread), tid: this thread
__device__ __forceinline__ bool goes_left(uint32_t w, const uint32_t *rk, int tid) {
  return rk[(w >> kSurRankBits) * kSurBlock + tid] <= (w & kSurRankMask);
}

// leaf index of the walk of one perfect tree of depth D (its blocks at bk)
template <int D>
__device__ __forceinline__ int walk(const uint4 *__restrict__ bk, const uint32_t *rk, int tid) {
  int p = 0, off = 0;
#pragma unroll
  for (int d = 0; d < D; d += 2) {
    const uint4 b = bk[off + p];
    const bool l0 = goes_left(b.x, rk, tid);
    p = 2 * p + (l0 ? 0 : 1);
    if (d + 1 < D) {
      const uint32_t w = l0 ? b.y : b.z;
      p = 2 * p + (goes_left(w, rk, tid) ? 0 : 1);
    }
    off += 1 << d;
  }
  return p;
}

// rank of x among n sorted thresholds t: #{t < x}; NaN above every index
__device__ __forceinline__ uint32_t rank_of(float x, const float *__restrict__ t, int n) {
  if (x != x) return 0xFFFFFFFFu;
  int lo = 0, hi = n;  // invariant: t[< lo] < x, t[>= hi] >= x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (t[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (uint32_t)lo;
}

template <int D>
__global__ void __launch_bounds__(kSurBlock)
surrogate_kernel(int64_t n, const double *__restrict__ nh, const double *__restrict__ tg,
                 const double *__restrict__ g0, const double *__restrict__ av,
                 const double *__restrict__ zeta, const double *__restrict__ fh2,
                 const double *__restrict__ fco, const SurrogateDev S, double xc_tot,
                 double xo_tot, const SurrogateOut out) {
#pragma clang fp contract(off)
  __shared__ uint32_t rk[kSurFeat * kSurBlock];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  if (i >= n) return;  // every access below is this thread's own
  constexpr int NB = block_count(D), NL = 1 << D;
  const double f[kSurFeat] = {log10(fmax(nh[i], 1.0)),    log10(fmax(tg[i], 5.0)),
                              log10(fmax(g0[i], 1e-2)),   av[i],
                              log10(fmax(zeta[i], 1e-18)), log10(fmax(fh2[i], 1e-10)),
                              log10(fmax(fco[i], 1e-10))};
#pragma unroll
  for (int k = 0; k < kSurFeat; ++k) {
    const float x = (float)((f[k] - S.x_mean[k]) / S.x_scale[k]);
    rk[k * kSurBlock + tid] = rank_of(x, S.thr + S.thr0[k], S.n_thr[k]);
  }
  double r[kSurOut];
#pragma unroll
  for (int o = 0; o < kSurOut; ++o) {
    double raw = S.base[o];
    const double sc = S.scale[o];
    const int nt = S.n_trees[o];
    const uint4 *nd = S.block + (int64_t)S.tree0[o] * NB;
    const double *lf = S.leaf + (int64_t)S.tree0[o] * NL;
    int t = 0;
    for (; t + 4 <= nt; t += 4) {
      const int a = walk<D>(nd + (int64_t)t * NB, rk, tid);
      const int b = walk<D>(nd + (int64_t)(t + 1) * NB, rk, tid);
      const int c = walk<D>(nd + (int64_t)(t + 2) * NB, rk, tid);
      const int e = walk<D>(nd + (int64_t)(t + 3) * NB, rk, tid);
      const double va = lf[(int64_t)t * NL + a], vb = lf[(int64_t)(t + 1) * NL + b];
      const double vc = lf[(int64_t)(t + 2) * NL + c], ve = lf[(int64_t)(t + 3) * NL + e];
      raw += sc * va;  // stage order of the library's accumulation
      raw += sc * vb;
      raw += sc * vc;
      raw += sc * ve;
    }
    for (; t < nt; ++t) raw += sc * lf[(int64_t)t * NL + walk<D>(nd + (int64_t)t * NB, rk, tid)];
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
  hipLaunchKernelGGL(surrogate_kernel<D>, dim3(grid_for(n, kSurBlock)), dim3(kSurBlock), 0, s,
                     n, in[0],
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
  // each feature's sorted distinct thresholds (equal values merged, so -0
  // and +0 share one rank) and every node's index into its feature's list
  std::vector<float> thr_all;
  std::vector<uint32_t> node_rank((size_t)(total * ni), 0u);
  for (int f = 0; f < kSurFeat; ++f) {
    std::vector<float> t;
    for (int64_t k = 0; k < total * ni; ++k)
      if (feature[k] == f) t.push_back(threshold[k]);
    for (float v : t) P3D_REQUIRE(v == v, "NaN threshold for feature %d", f);
    std::sort(t.begin(), t.end());
    t.erase(std::unique(t.begin(), t.end(), [](float a, float b) { return a == b; }), t.end());
    P3D_REQUIRE(t.size() <= (size_t)kSurRankMask, "too many distinct thresholds for feature %d",
                f);
    d.thr0[f] = (int)thr_all.size();
    d.n_thr[f] = (int)t.size();
    for (int64_t k = 0; k < total * ni; ++k)
      if (feature[k] == f)
        node_rank[k] = (uint32_t)(std::lower_bound(t.begin(), t.end(), threshold[k]) - t.begin());
    thr_all.insert(thr_all.end(), t.begin(), t.end());
  }
  const int nb = block_count(depth);
  const size_t blk_bytes = (size_t)(total * nb) * sizeof(uint4);
  const size_t leaf_off = (blk_bytes + 255) & ~(size_t)255;
  const size_t thr_off = (leaf_off + (size_t)(total * nl) * sizeof(double) + 255) & ~(size_t)255;
  const size_t bytes = thr_off + std::max<size_t>(thr_all.size(), 1) * sizeof(float);
  std::vector<char> host(bytes, 0);
  uint4 *hb = reinterpret_cast<uint4 *>(host.data());
  auto word = [&](int64_t t, int64_t node) -> uint32_t {
    const int64_t k = t * ni + node;
    return ((uint32_t)feature[k] << kSurRankBits) | node_rank[k];
  };
  for (int64_t t = 0; t < total; ++t) {
    int off = 0;
    for (int dd = 0; dd < depth; dd += 2) {
      for (int p = 0; p < (1 << dd); ++p) {
        const int64_t node = ((int64_t)1 << dd) - 1 + p;  // breadth-first index
        uint4 b = make_uint4(word(t, node), 0u, 0u, 0u);
        if (dd + 1 < depth) {
          b.y = word(t, 2 * node + 1);
          b.z = word(t, 2 * node + 2);
        }
        hb[t * nb + off + p] = b;
      }
      off += 1 << dd;
    }
  }
  std::memcpy(host.data() + leaf_off, leaf, (size_t)(total * nl) * sizeof(double));
  if (!thr_all.empty())
    std::memcpy(host.data() + thr_off, thr_all.data(), thr_all.size() * sizeof(float));
  if (bytes > 0) {
    if (hipMalloc(&sur->blob, bytes) != hipSuccess) {
      set_error("hipMalloc of %zu bytes for the tree ensemble failed", bytes);
      return P3D_ENOMEM;
    }
    P3D_CHECK_HIP(hipMemcpy(sur->blob, host.data(), bytes, hipMemcpyHostToDevice));
  }
  d.block = reinterpret_cast<const uint4 *>(sur->blob);
  d.leaf = reinterpret_cast<const double *>(static_cast<char *>(sur->blob) + leaf_off);
  d.thr = reinterpret_cast<const float *>(static_cast<char *>(sur->blob) + thr_off);
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
  if (sur->mlp) {
    const int64_t blocks = (n + kMlpRows - 1) / kMlpRows;
    P3D_REQUIRE(blocks <= 0x7fffffffLL, "too many cells for one launch");
    hipLaunchKernelGGL(surrogate_mlp_kernel, dim3((unsigned)blocks), dim3(64 * kMlpWaves), 0,
                       as_stream(stream), n, in[0], in[1], in[2], in[3], in[4], in[5], in[6],
                       sur->d, sur->m, x_c_total, x_o_total, o);
    return launch_status("surrogate_mlp_kernel");
  }
  return launch_surrogate(sur->depth, as_stream(stream), n, in, sur->d, x_c_total, x_o_total, o);
}

int p3d_surrogate_mlp_create(int32_t n_outputs, int32_t n_features, int32_t n_hidden,
                             const int32_t *hidden, const double *weights, const double *x_mean,
                             const double *x_scale, const double *y_mean, const double *y_scale,
                             int64_t *handle) {
  P3D_REQUIRE(handle, "handle pointer is required");
  P3D_REQUIRE(n_outputs == kSurOut && n_features == kSurFeat,
              "the surrogate maps %d features to %d species (got %d -> %d)", kSurFeat, kSurOut,
              n_features, n_outputs);
  P3D_REQUIRE(n_hidden >= 1 && n_hidden <= kMlpMaxLayers - 1,
              "n_hidden must be in [1, %d] (got %d)", kMlpMaxLayers - 1, n_hidden);
  P3D_REQUIRE(hidden && weights && x_mean && x_scale && y_mean && y_scale, "null host array");
  int width[kMlpMaxLayers + 1];
  width[0] = kSurFeat;
  for (int l = 0; l < n_hidden; ++l) {
    P3D_REQUIRE(hidden[l] >= 1 && hidden[l] <= kMlpMaxWidth,
                "hidden layer %d width must be in [1, %d] (got %d)", l, kMlpMaxWidth, hidden[l]);
    width[l + 1] = hidden[l];
  }
  width[n_hidden + 1] = 1;
  auto sur = std::make_unique<Surrogate>();
  sur->mlp = true;
  MlpDev &m = sur->m;
  m.n_layers = n_hidden + 1;
  int64_t blk = 0;
  for (int l = 0; l < m.n_layers; ++l) {
    m.kpad[l] = l == 0 ? 8 : (width[l] + 15) / 16 * 16;  // inputs: the previous padded width
    m.npad[l] = (width[l + 1] + 15) / 16 * 16;
    m.woff[l] = blk;
    blk += (int64_t)m.kpad[l] * m.npad[l];
    m.boff[l] = blk;
    blk += m.npad[l];
  }
  m.blk = blk;
  // the caller's weights: per species, per layer W (in x out, row-major)
  // then b (out), unpadded; laid out padded here
  std::vector<double> host((size_t)(blk * kSurOut), 0.0);
  const double *src = weights;
  for (int o = 0; o < kSurOut; ++o) {
    double *dst = host.data() + (size_t)o * blk;
    for (int l = 0; l < m.n_layers; ++l) {
      const int in = width[l], outw = width[l + 1];
      for (int k = 0; k < in; ++k)
        for (int j = 0; j < outw; ++j) dst[m.woff[l] + (int64_t)k * m.npad[l] + j] = *src++;
      for (int j = 0; j < outw; ++j) dst[m.boff[l] + j] = *src++;
    }
  }
  for (size_t k = 0; k < host.size(); ++k)
    P3D_REQUIRE(std::isfinite(host[k]), "non-finite MLP weight");
  SurrogateDev &d = sur->d;
  for (int o = 0; o < kSurOut; ++o) {
    d.y_mean[o] = y_mean[o];
    d.y_scale[o] = y_scale[o];
  }
  for (int k = 0; k < kSurFeat; ++k) {
    d.x_mean[k] = x_mean[k];
    d.x_scale[k] = x_scale[k];
  }
  const size_t bytes = host.size() * sizeof(double);
  if (hipMalloc(&sur->blob, bytes) != hipSuccess) {
    set_error("hipMalloc of %zu bytes for the MLP weights failed", bytes);
    return P3D_ENOMEM;
  }
  P3D_CHECK_HIP(hipMemcpy(sur->blob, host.data(), bytes, hipMemcpyHostToDevice));
  m.w = static_cast<const double *>(sur->blob);
  std::lock_guard<std::mutex> lk(g_sur_mu);
  const int64_t h = g_next_sur++;
  g_surs[h] = std::move(sur);
  *handle = h;
  return P3D_OK;
}

}  // extern "C"
