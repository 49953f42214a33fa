// Species network on gfx950: rate coefficients, stiff RHS, analytic
// Jacobian and the implicit variable-order BDF solve of every cell.
//
// Mapping: one cell per 64-lane workgroup (one wavefront).  Lane s owns
// species s (its abundance and its backward differences live in registers);
// lane r also evaluates reactions r, r+64, ...  The Jacobian and the
// iteration matrix M = I - c J (row-major, padded stride) are staged in LDS;
// the LU pivot search is a wavefront argmax, the triangular solves keep the
// right-hand side in registers and broadcast with readlane, and every norm is
// a butterfly reduction.  All per-wave control values (step size, order,
// counters) are wave-uniform, so the whole solve runs with scalar branches.
//
// Upstream behaviour restated (paths relative to the upstream repository):
//   rates + RHS            prism3d/chemistry/solver.py:69-143
//   steady-state driver    prism3d/chemistry/solver.py:145-225
//   H+/e- correction       prism3d/chemistry/solver.py:227-281
//   conservation           prism3d/chemistry/solver.py:345-388
//   integrator             scipy solve_ivp(method='BDF') as called at
//                          solver.py:195-201: the quasi-constant step size NDF
//                          of Shampine & Reichelt (1997), orders 1-5.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "lane_launch.h"
#include "p3d_common.h"
#include "wave_net.h"

namespace p3d {

constexpr int kMaxSpecies = 64;       // one-wavefront forms: lane = species
constexpr int kMaxSpeciesWide = 128;  // two-wavefront form of the solve
constexpr int kMaxPieces = kMaxSpeciesWide + 128;  // RHS pieces: rows + lane cuts
constexpr int kRxPerLaneCap = 2;  // reactions per lane of the register cache (kRxPerLane)
constexpr int kMaxOrder = 5;
constexpr int kNumD = kMaxOrder + 3;

struct NetDev {
  int ns, nr, e_idx, nnz, n_rhs_e;
  int roles[7];  // H, H2, H+, C+, C, CO, O
  const int *rtype;
  const double *alpha, *beta, *gamma;
  const int *react;   // 2 per reaction
  const int *dep;     // 2 per reaction: species the rate depends on (or -1)
  const int *st5;     // 5 per reaction: consumed species (2 slots) then produced (3), or -1
  // per species row of the Jacobian (networks of <= 32 species): its first
  // structural entry in the jrow/jcol order and the mask of its structural
  // columns (the register-LU form keeps J sparse)
  const int *jr_start;
  const unsigned *jr_mask;
  const int *rhs_off; // ns+1
  const int *rhs_rxn;
  const double *rhs_sgn;
  const int *rhs_pk;  // packed RHS entries: (reaction << 1) | (product ? 1 : 0)
  const int *jrow, *jcol, *joff;  // nnz, nnz, nnz+1
  const int *jt_rxn, *jt_slot;
  const double *jt_coef;
  // the same Jacobian sums on a balanced schedule (one-wavefront register-LU
  // form): whole entries spread over the 64 lanes longest first, each summed
  // in its own term order; lane l's slot s is jsched[64 s + l], packed as
  // part index | int8 coefficient << 8 | entry << 16 | last term << 25 |
  // used << 26 (0 when the network does not fit the packing)
  int jslots;
  const unsigned *jsched;
  const double *post_w;  // 4 * ns
  // balanced RHS gather (the two-wavefront form, and the one-wavefront BDF
  // solve of networks past the register cache): the packed entries (rhs_pk,
  // species by species) cut into pieces at every species boundary and at 128
  // equal shares.  The exported RHS, polish and steady adjoint sum each row
  // serially (rhs_lane); a row split over shares differs from it by rounding.
  int n_pc;
  const int *pc_beg;   // n_pc + 1 entry offsets
  const int *pc_lane;  // 129: pieces of lane l are pc_lane[l] .. pc_lane[l+1]-1
  const int *pc_row;   // ns + 1: pieces of species s
  // static elimination schedule of the sparse no-pivot LU (LinSparseG), as
  // p3d_sparse_lu_schedule returns it: species s sits at position sl_pos[s]
  // of the minimum-Markowitz order; sl_lev[p] = dependency level of position
  // p (read by the factorisation); the pivots of level v are sl_lvp[
  // sl_lvoff[v] .. sl_lvoff[v+1]-1] (position order); row position p of the
  // filled structure has upper entries at sl_uj[sl_uoff[p] .. sl_uoff[p+1]-1]
  // (the kernels walk the packed LDS copy below instead)
  int sl_nlev;
  const int *sl_pos, *sl_lev, *sl_lvoff, *sl_lvp, *sl_uoff, *sl_uj;
  // the same schedule packed for LDS (LinSparseG): descriptor q of the
  // pivots in level order = k | len << 7 | ubeg << 14 | last of its level
  // << 31, len upper entries at bytes ubeg.. of sl_ub (each list starts on
  // a 4-byte boundary); sl_zb[z] = position of column jcol[z]; sl_pb[s] =
  // position of species s.  Byte arrays padded to whole words.
  int sl_nub, sl_nzb;
  const unsigned *sl_desc, *sl_ub, *sl_zb, *sl_pb;
};

struct Network {
  NetDev d;
  void *blob = nullptr;  // single device allocation holding all tables
  bool lane_ok = false;  // structure matches the generated one-cell-per-lane code
  // slot memory of the persistent lane kernel (grown on demand) and its
  // work counter; calls on one handle share it, so they run in stream order
  void *lane_ws = nullptr;
  size_t lane_ws_bytes = 0;
  // work counter of the persistent one-cell-per-wavefront kernels
  unsigned long long *queue_ctr = nullptr;
  // per resident wavefront of the tape adjoint: one segment's stored factors
  void *segbuf = nullptr;
  size_t seg_bytes = 0;
  ~Network() {
    if (blob) (void)hipFree(blob);
    if (lane_ws) (void)hipFree(lane_ws);
    if (queue_ctr) (void)hipFree(queue_ctr);
    if (segbuf) (void)hipFree(segbuf);
  }
};

static std::mutex g_net_mu;
static std::map<int64_t, std::unique_ptr<Network>> g_nets;
static int64_t g_next_handle = 1;

static Network *find_net(int64_t h) {
  std::lock_guard<std::mutex> lk(g_net_mu);
  auto it = g_nets.find(h);
  return it == g_nets.end() ? nullptr : it->second.get();
}

struct CellParams {
  double nH, T, G0, AV, zeta, fH2, fCO;
  double g0_sqrt_t;  // G0 sqrt(T) of the PAH charge balance, formed once per cell
};

// LDS layout of one cell (one workgroup), all doubles unless noted.
struct Lds {
  double *kco;   // nr rate coefficients
  double *part;  // 2*nr partial derivatives (also the rates)
  double *ys;    // 64 floored abundances
  double *J;     // ns x ld
  double *M;     // ns x ld
  int *piv;      // 64
  double *ru;    // 36: step-size change matrix RU of the backward differences
  int *csr_off;  // 65: per-species offsets into csr (LDS copy of rhs_off)
  int *csr;      // n_rhs_e packed RHS entries (LDS copy of rhs_pk)
  double *red;   // two-wavefront form: 4 doubles of cross-wave reduction slots
  double *xs;    // components of the triangular solves (128 two-wavefront, 32 register LU)
  double *psum;  // two-wavefront form: partial sums of the RHS pieces
  // static sparse LU (LinSparseG), staged from NetDev once per workgroup:
  // one descriptor per pivot, the upper-row positions and the scatter
  // positions of the Jacobian entries as bytes, the species positions
  unsigned *sd;
  unsigned char *su, *sz, *sp;
  int ld;
};

// with_m = false: no dense M (linear-algebra forms that keep M in registers);
// wave_pieces: the 64 piece sums of rhs_pieces.
__device__ __forceinline__ Lds carve(double *base, const NetDev &N, bool with_m = true,
                                     bool wave_pieces = false) {
  Lds L;
  L.ld = N.ns | 1;
  L.kco = base;
  L.part = L.kco + N.nr;
  L.ys = L.part + 2 * N.nr;
  L.J = L.ys + kMaxSpecies;
  L.M = with_m ? L.J + N.ns * L.ld : nullptr;
  L.piv = reinterpret_cast<int *>(L.J + (with_m ? 2 : 1) * N.ns * L.ld);
  L.ru = reinterpret_cast<double *>(L.piv + kMaxSpecies);
  // piece sums of the balanced RHS, for networks past the register cache,
  // and of the register-LU form's RHS pieces (one per lane)
  const bool pieces = N.nr > 64 * kRxPerLaneCap;
  L.psum = (pieces || wave_pieces) ? L.ru + 36 : nullptr;
  // the register-LU form's solves record their 32 components in xs
  L.xs = wave_pieces ? L.ru + 36 + 64 : nullptr;
  L.csr_off = reinterpret_cast<int *>(L.ru + 36 +
                                      (pieces ? kMaxPieces : (wave_pieces ? 64 + 32 : 0)));
  L.csr = L.csr_off + kMaxSpecies + 1;
  L.red = nullptr;
  return L;
}

// LDS of one cell in the two-wavefront form (up to 128 species): J is kept
// sparse (its nnz structural entries, in the order of jrow/jcol) beside the
// dense M, and the RHS lists are read from global memory (L2).
__device__ __forceinline__ Lds carve_wide(double *base, const NetDev &N) {
  Lds L;
  L.ld = N.ns | 1;
  L.kco = base;
  L.part = L.kco + N.nr;
  L.ys = L.part + 2 * N.nr;
  L.J = L.ys + kMaxSpeciesWide;
  L.M = L.J + N.nnz;
  L.red = L.M + N.ns * L.ld;
  L.xs = L.red + 4;
  L.psum = L.xs + kMaxSpeciesWide;
  L.ru = L.psum + kMaxPieces;
  L.piv = reinterpret_cast<int *>(L.ru + 36);
  L.csr_off = nullptr;
  L.csr = nullptr;
  L.sd = reinterpret_cast<unsigned *>(L.piv + kMaxSpeciesWide);
  L.su = reinterpret_cast<unsigned char *>(L.sd + kMaxSpeciesWide);
  L.sz = L.su + N.sl_nub;
  L.sp = L.sz + N.sl_nzb;
  return L;
}

// LDS of one cell in the register-LU form (bundled network): J kept sparse
// (its nnz structural entries in the jrow/jcol order), M dense for the
// factors, the step-size change matrix, the 64 RHS piece sums and the 32
// components of the solves.  No CSR lists (each lane's RHS piece is in
// registers) and no pivots.
// The sizes are the generated network's (wave_net.h), so every LDS address
// is the array base plus a constant: no pointer is held in registers.
__device__ __forceinline__ Lds carve_wave(double *base, const NetDev &) {
  constexpr int ld = wave::kNS | 1;
  Lds L;
  L.ld = ld;
  L.kco = base;
  L.part = base + wave::kNR;
  L.ys = base + 3 * wave::kNR;
  L.J = base + 3 * wave::kNR + kMaxSpecies;
  L.M = base + 3 * wave::kNR + kMaxSpecies + wave::kNNZ + 1;  // J's spare slot
  L.ru = L.M + (wave::kNS + 1) * ld;
  L.psum = L.ru + 36;
  L.xs = L.psum + 64;
  L.piv = nullptr;
  L.csr_off = nullptr;
  L.csr = nullptr;
  L.red = nullptr;
  return L;
}
static size_t lds_bytes_wave() {
  return sizeof(double) * (size_t)(3 * wave::kNR + kMaxSpecies + wave::kNNZ + 1 +
                                   (wave::kNS + 1) * (wave::kNS | 1) + 36 + 64 + 32);
}

static size_t lds_bytes(int ns, int nr, int n_rhs_e, bool with_m = true,
                        bool wave_pieces = false) {
  int ld = ns | 1;
  const int pieces = nr > 64 * kRxPerLaneCap ? kMaxPieces : 0;
  return sizeof(double) * (size_t)(3 * nr + kMaxSpecies + (with_m ? 2 : 1) * ns * ld + 36 +
                                   (pieces ? pieces : (wave_pieces ? 64 + 32 : 0))) +
         sizeof(int) * (size_t)(2 * kMaxSpecies + 1 + n_rhs_e);
}

static size_t lds_bytes_wide(int ns, int nr, int nnz) {
  const int ld = ns | 1;
  return sizeof(double) * (size_t)(3 * nr + 2 * kMaxSpeciesWide + kMaxPieces + nnz + ns * ld +
                                   4 + 36) +
         sizeof(int) * (size_t)kMaxSpeciesWide;
}
// the two-wavefront form with the static sparse LU's staged schedule
static size_t lds_bytes_wide_sparse(const NetDev &d) {
  return lds_bytes_wide(d.ns, d.nr, d.nnz) + sizeof(unsigned) * (size_t)kMaxSpeciesWide +
         (size_t)d.sl_nub + (size_t)d.sl_nzb + (size_t)kMaxSpeciesWide;
}

// Workgroup collectives.  Wave1: one wavefront, the DPP/readlane collectives.
// Wave2: two wavefronts (128 lanes); each wavefront reduces on its own and
// the two results combine through LDS in a fixed order (wavefront 0 first),
// so every lane of both gets the same bits and branches on them stay
// uniform across the workgroup.  Every lane of the workgroup must call.
struct Wave1 {
  static constexpr int W = 64;
  double *red;
  __device__ __forceinline__ double sum(double v) const { return wave_sum(v); }
  // every lane's v finite (the per-element test of the reference's Newton
  // iteration), as one ballot
  __device__ __forceinline__ bool all_finite(double v) const {
    return __builtin_amdgcn_ballot_w64(!isfinite(v)) == 0;
  }
  __device__ __forceinline__ double bc(double v, int src) const { return bcast(v, src); }
  __device__ __forceinline__ int argmax(double v, int lane) const {
    return wave_argmax(v, lane);
  }
};
struct Wave2 {
  static constexpr int W = 128;
  double *red;  // 4 doubles of LDS
  __device__ __forceinline__ double sum(double v) const {
    const double s = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    const double r = red[0] + red[1];
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ bool all_finite(double v) const {
    return sum(isfinite(v) ? 0.0 : 1.0) == 0.0;
  }
  __device__ __forceinline__ double bc(double v, int src) const {
    if ((int)threadIdx.x == src) red[0] = v;
    __syncthreads();
    const double r = red[0];
    __syncthreads();
    return r;
  }
  // largest v and the lowest lane holding it
  __device__ __forceinline__ int argmax(double v, int lane) const {
    const int i = wave_argmax(v, lane);
    const double bv = bcast(v, i & 63);
    int *ri = reinterpret_cast<int *>(red + 2);
    if ((lane & 63) == 0) {
      red[lane >> 6] = bv;
      ri[lane >> 6] = i;
    }
    __syncthreads();
    const double v0 = red[0], v1 = red[1];
    const int i0 = ri[0], i1 = ri[1];
    __syncthreads();
    return uniform(v1 > v0 ? i1 : i0);  // on a tie wavefront 0 holds the lower lane
  }
};

template <class G>
__device__ __forceinline__ double rms_g(const G &g, double v, int ns) {
  return sqrt(g.sum(v * v) / ns);
}

// Copies the RHS stoichiometry lists into LDS (once per workgroup).
__device__ __forceinline__ void stage_tables(const NetDev &N, const Lds &L, int lane) {
  for (int i = lane; i <= N.ns; i += 64) L.csr_off[i] = N.rhs_off[i];
  for (int i = lane; i < N.n_rhs_e; i += 64) L.csr[i] = N.rhs_pk[i];
}

// Reactions evaluated by one lane (r = lane + 64 s), kept in registers.
constexpr int kRxPerLane = kRxPerLaneCap;
struct RxCache {
  int t[kRxPerLane], a[kRxPerLane], b[kRxPerLane];
  double k[kRxPerLane];
  // the rate's last factor by type (n_H two-body, n_H / 2 grain, 1 else), so
  // rate = ((k y_a) y_b) fac with y_b = 1 when there is none: the same
  // rounding as the per-type products (halving is exact)
  __device__ __forceinline__ double fac(int q, const CellParams &p) const {
    return t[q] == P3D_RX_TWO_BODY ? p.nH : (t[q] == P3D_RX_GRAIN ? p.nH * 0.5 : 1.0);
  }
  __device__ __forceinline__ void load(const NetDev &N, const Lds &L, int lane) {
#pragma unroll
    for (int q = 0; q < kRxPerLane; ++q) {
      const int r = lane + 64 * q;
      const bool v = r < N.nr;
      t[q] = v ? N.rtype[r] : -1;
      a[q] = v ? N.react[2 * r] : 0;
      b[q] = v ? N.react[2 * r + 1] : 0;
      k[q] = v ? L.kco[r] : 0.0;
    }
  }
};

// Rate coefficient with its exponentials in FP32: one expf of the combined
// exponent per reaction, the prefactors in FP64 (relative error about
// |exponent| * 2^-24 against the FP64 form).
__device__ __forceinline__ double rate_coeff_f32(int t, double a, double b, double g,
                                                 const CellParams &p) {
  const float Tc = fminf(fmaxf((float)p.T, 10.0f), 41000.0f);
  const float bf = (float)b, gf = (float)g;
  switch (t) {
    case P3D_RX_TWO_BODY: return a * (double)expf(bf * logf(Tc / 300.0f) - gf / Tc);
    case P3D_RX_PHOTO: return a * p.G0 * (double)expf(-gf * (float)p.AV);
    case P3D_RX_PHOTO_FH2: return a * p.G0 * (double)expf(-gf * (float)p.AV) * p.fH2;
    case P3D_RX_PHOTO_FCO: return a * p.G0 * (double)expf(-gf * (float)p.AV) * p.fCO;
    case P3D_RX_CR: return a * p.zeta;
    default: return a * (double)expf(bf * logf((float)p.T / 300.0f));  // grain surface
  }
}

// Rate coefficients k_r (solver.py:100-133); f32: exponentials in FP32.
// W: lanes of the workgroup.
template <int W = 64>
__device__ void rate_coeffs(const NetDev &N, const CellParams &p, const Lds &L, int lane,
                            bool f32 = false) {
  const double Tc = fmin(fmax(p.T, 10.0), 41000.0);
  const double lTc = log(Tc / 300.0), lT = log(p.T / 300.0);
  for (int r = lane; r < N.nr; r += W) {
    const double a = N.alpha[r], b = N.beta[r], g = N.gamma[r];
    if (f32) {
      L.kco[r] = rate_coeff_f32(N.rtype[r], a, b, g, p);
      continue;
    }
    double k;
    switch (N.rtype[r]) {
      case P3D_RX_TWO_BODY: k = a * exp(b * lTc) * exp(-g / Tc); break;
      case P3D_RX_PHOTO: k = a * p.G0 * exp(-g * p.AV); break;
      case P3D_RX_PHOTO_FH2: k = a * p.G0 * exp(-g * p.AV) * p.fH2; break;
      case P3D_RX_PHOTO_FCO: k = a * p.G0 * exp(-g * p.AV) * p.fCO; break;
      case P3D_RX_CR: k = a * p.zeta; break;
      default: k = a * exp(b * lT); break;  // grain surface
    }
    L.kco[r] = k;
  }
}

// PAH- abundance from the electron abundance (solver.py:74-87) and its
// derivative with respect to x_e.
__device__ __forceinline__ void pah_minus(double ye_floored, const CellParams &p,
                                          double &pm, double &dpm) {
  const double xpah = 1e-7;
  double ye = fmax(ye_floored, 1e-10);
  double ne = ye * p.nH;
  if (p.G0 > 0.0 && ne > 0.0) {
    double psi = p.g0_sqrt_t / ne;
    double fm = fmax(500.0 / (psi + 500.0) * 0.5, 0.0);
    pm = xpah * fm;
    dpm = ye_floored > 1e-10 ? xpah * 250.0 * psi / ((psi + 500.0) * (psi + 500.0) * ye) : 0.0;
  } else {
    pm = xpah * 0.4;
    dpm = 0.0;
  }
}

// Publishes max(y, 1e-30) of lane s into ys[s].
__device__ __forceinline__ void publish_y(const NetDev &N, const Lds &L, double y, int lane) {
  if (lane < N.ns) L.ys[lane] = fmax(y, 1e-30);
  __syncthreads();
}

// Reaction rates into part[2r] (want_jac: partials into part[2r], part[2r+1]
// instead).  Requires ys published.  W: lanes of the workgroup.
template <int W = 64>
__device__ void eval_reactions(const NetDev &N, const CellParams &p, const Lds &L, int lane,
                               bool want_jac) {
  double pm, dpm;
  pah_minus(L.ys[N.e_idx], p, pm, dpm);
  for (int r = lane; r < N.nr; r += W) {
    const double k = L.kco[r];
    const int t = N.rtype[r];
    const int a = N.react[2 * r], b = N.react[2 * r + 1];
    const double ya = L.ys[a];
    if (!want_jac) {
      double rate;
      if (t == P3D_RX_TWO_BODY) {
        double yb = (b == P3D_SP_PAH_MINUS) ? pm : L.ys[b];
        rate = k * ya * yb * p.nH;
      } else if (t == P3D_RX_GRAIN) {
        rate = k * ya * p.nH * 0.5;
      } else {
        rate = k * ya;
      }
      L.part[2 * r] = rate;
    } else {
      // partials with respect to dep[2r], dep[2r+1]; zero below the floor
      const double act_a = ya > 1e-30 ? 1.0 : 0.0;
      double d0 = 0.0, d1 = 0.0;
      if (t == P3D_RX_TWO_BODY) {
        if (b == P3D_SP_PAH_MINUS) {
          d0 = k * pm * p.nH * act_a;
          d1 = k * ya * p.nH * dpm * (L.ys[N.e_idx] > 1e-30 ? 1.0 : 0.0);
        } else if (a == b) {
          d0 = 2.0 * k * ya * p.nH * act_a;
        } else {
          const double yb = L.ys[b];
          d0 = k * yb * p.nH * act_a;
          d1 = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0);
        }
      } else if (t == P3D_RX_GRAIN) {
        d0 = k * p.nH * 0.5 * act_a;
      } else {
        d0 = k * act_a;
      }
      L.part[2 * r] = d0;
      L.part[2 * r + 1] = d1;
    }
  }
  __syncthreads();
}

// dy/dt of species `lane` from the register reaction cache and the LDS
// stoichiometry lists (networks with <= 128 reactions).  Same accumulation
// order as rhs_lane.  Requires y published.
__device__ __forceinline__ double rhs_cached(const NetDev &N, const CellParams &p, const Lds &L,
                                             const RxCache &rc, int lane) {
  double pm, dpm;
  pah_minus(L.ys[N.e_idx], p, pm, dpm);
#pragma unroll
  for (int q = 0; q < kRxPerLane; ++q) {
    const int t = rc.t[q];
    if (t < 0) continue;
    const double ya = L.ys[rc.a[q]];
    double rate;
    if (t == P3D_RX_TWO_BODY) {
      const double yb = (rc.b[q] == P3D_SP_PAH_MINUS) ? pm : L.ys[rc.b[q]];
      rate = rc.k[q] * ya * yb * p.nH;
    } else if (t == P3D_RX_GRAIN) {
      rate = rc.k[q] * ya * p.nH * 0.5;
    } else {
      rate = rc.k[q] * ya;
    }
    L.part[2 * (lane + 64 * q)] = rate;
  }
  __syncthreads();
  double acc = 0.0;
  if (lane < N.ns) {
    const int e0 = L.csr_off[lane], e1 = L.csr_off[lane + 1];
    int e = e0;
    for (; e + 4 <= e1; e += 4) {
      const int c0 = L.csr[e], c1 = L.csr[e + 1], c2 = L.csr[e + 2], c3 = L.csr[e + 3];
      const double v0 = L.part[2 * (c0 >> 1)], v1 = L.part[2 * (c1 >> 1)],
                   v2 = L.part[2 * (c2 >> 1)], v3 = L.part[2 * (c3 >> 1)];
      acc = (c0 & 1) ? acc + v0 : acc - v0;
      acc = (c1 & 1) ? acc + v1 : acc - v1;
      acc = (c2 & 1) ? acc + v2 : acc - v2;
      acc = (c3 & 1) ? acc + v3 : acc - v3;
    }
    for (; e < e1; ++e) {
      const int c0 = L.csr[e];
      const double v0 = L.part[2 * (c0 >> 1)];
      acc = (c0 & 1) ? acc + v0 : acc - v0;
    }
  }
  __syncthreads();  // part[] is reused by the next evaluation
  return acc;
}

// RHS pieces of the register-LU form (bundled network, wave_net.h): lane l
// sums piece l (at most kPieceLen entries of one species' row, in the
// upstream order, their indices held in two registers), then species s adds
// its pieces in order.  A row in one piece gets the bits of rhs_cached.
__constant__ unsigned c_wave_pw0[64] = P3D_WAVE_PIECE_W0;
__constant__ unsigned c_wave_pw1[64] = P3D_WAVE_PIECE_W1;
__constant__ int c_wave_rowpc[wave::kNS + 1] = P3D_WAVE_ROW_PIECE;

__device__ __forceinline__ double rhs_pieces(const NetDev &N, const CellParams &p, const Lds &L,
                                             const RxCache &rc, unsigned pw0, unsigned pw1,
                                             int lane) {
  double pm, dpm;
  pah_minus(L.ys[N.e_idx], p, pm, dpm);
  // branch-free rates: every lane evaluates the same products
#pragma unroll
  for (int q = 0; q < kRxPerLane; ++q) {
    const int t = rc.t[q];
    const double ya = L.ys[rc.a[q]];
    const int b = rc.b[q];
    const bool two = t == P3D_RX_TWO_BODY;
    const double yb = !two ? 1.0 : (b == P3D_SP_PAH_MINUS ? pm : L.ys[b < 0 ? 0 : b]);
    const double rate = ((rc.k[q] * ya) * yb) * rc.fac(q, p);
    // every lane holds a reaction in the first round (64 < kNR): no branch
    if (64 * q + 63 < wave::kNR || t >= 0) L.part[2 * (lane + 64 * q)] = rate;
  }
  __syncthreads();
  // piece sums without branches: an unused slot reads part[0] and keeps acc;
  // acc + (-v) is acc - v bit for bit
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < wave::kPieceLen; ++q) {
    const unsigned e = ((q < 4 ? pw0 : pw1) >> (8 * (q & 3))) & 0xFFu;
    const bool use = e != 0xFFu;
    const double v = L.part[use ? 2 * (e >> 1) : 0];
    const double sv = (e & 1u) ? v : -v;
    acc = use ? acc + sv : acc;
  }
  L.psum[lane] = acc;
  __syncthreads();
  // the species' pieces in order, as a fixed-length predicated sum so the
  // reads are issued together (slots past the row keep f by a select)
  double f = 0.0;
  {
    const int q0 = lane < N.ns ? c_wave_rowpc[lane] : 0;
    const int q1 = lane < N.ns ? c_wave_rowpc[lane + 1] : 0;
#pragma unroll
    for (int j = 0; j < wave::kRowPieceMax; ++j) {
      const int q = q0 + j;
      const double v = L.psum[q < 64 ? q : 63];
      f = q < q1 ? f + v : f;
    }
  }
  __syncthreads();  // part[] and psum[] are reused by the next evaluation
  return f;
}

// dy/dt of species `lane` (upstream accumulation order).  Requires y published.
template <int W = 64>
__device__ double rhs_lane(const NetDev &N, const CellParams &p, const Lds &L, int lane) {
  eval_reactions<W>(N, p, L, lane, false);
  double acc = 0.0;
  if (lane < N.ns) {
    const int e0 = N.rhs_off[lane], e1 = N.rhs_off[lane + 1];
    for (int e = e0; e < e1; ++e) acc += N.rhs_sgn[e] * L.part[2 * N.rhs_rxn[e]];
  }
  __syncthreads();  // part[] is reused by the next evaluation
  return acc;
}

// Analytic Jacobian into L.J (row-major, stride ld).  Requires y published.
__device__ void jacobian(const NetDev &N, const CellParams &p, const Lds &L, int lane) {
  eval_reactions(N, p, L, lane, true);
  for (int idx = lane; idx < N.ns * L.ld; idx += 64) L.J[idx] = 0.0;
  __syncthreads();
  for (int z = lane; z < N.nnz; z += 64) {
    double acc = 0.0;
    for (int t = N.joff[z]; t < N.joff[z + 1]; ++t)
      acc += N.jt_coef[t] * L.part[2 * N.jt_rxn[t] + N.jt_slot[t]];
    L.J[N.jrow[z] * L.ld + N.jcol[z]] = acc;
  }
  __syncthreads();
}

// The Jacobian's structural entries into L.J[z] (z in the order of
// jrow/jcol), the same sums as the dense form.  Requires y published.
template <int W>
__device__ void jacobian_sparse(const NetDev &N, const CellParams &p, const Lds &L, int lane) {
  eval_reactions<W>(N, p, L, lane, true);
  for (int z = lane; z < N.nnz; z += W) {
    double acc = 0.0;
    for (int t = N.joff[z]; t < N.joff[z + 1]; ++t)
      acc += N.jt_coef[t] * L.part[2 * N.jt_rxn[t] + N.jt_slot[t]];
    L.J[z] = acc;
  }
  __syncthreads();
}

// jacobian_sparse's sums on the balanced schedule (N.jsched): every lane runs
// the same N.jslots slots, so no lane waits on the longest entry's loop and
// the slots' table reads are independent of the data.  Each entry is summed
// by one lane in its term order from 0, bit for bit as jacobian_sparse; a
// slot that ends no entry writes the spare slot J[kNNZ].
// eval_reactions' partials (want_jac) from the lane's register reaction
// cache (reaction lane + 64 q), the same expressions in the same order
__device__ __forceinline__ void partials_cached(const NetDev &N, const CellParams &p,
                                                const Lds &L, const RxCache &rc, int lane) {
  double pm, dpm;
  pah_minus(L.ys[N.e_idx], p, pm, dpm);
#pragma unroll
  for (int q = 0; q < kRxPerLane; ++q) {
    const int t = rc.t[q];
    if (t < 0) continue;
    const int a = rc.a[q], b = rc.b[q];
    const double k = rc.k[q];
    const double ya = L.ys[a];
    const double act_a = ya > 1e-30 ? 1.0 : 0.0;
    double d0 = 0.0, d1 = 0.0;
    if (t == P3D_RX_TWO_BODY) {
      if (b == P3D_SP_PAH_MINUS) {
        d0 = k * pm * p.nH * act_a;
        d1 = k * ya * p.nH * dpm * (L.ys[N.e_idx] > 1e-30 ? 1.0 : 0.0);
      } else if (a == b) {
        d0 = 2.0 * k * ya * p.nH * act_a;
      } else {
        const double yb = L.ys[b];
        d0 = k * yb * p.nH * act_a;
        d1 = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0);
      }
    } else if (t == P3D_RX_GRAIN) {
      d0 = k * p.nH * 0.5 * act_a;
    } else {
      d0 = k * act_a;
    }
    L.part[2 * (lane + 64 * q)] = d0;
    L.part[2 * (lane + 64 * q) + 1] = d1;
  }
  __syncthreads();
}

template <int W>
__device__ void jacobian_sched(const NetDev &N, const CellParams &p, const Lds &L, int lane,
                               const RxCache *rc = nullptr) {
  if (rc) partials_cached(N, p, L, *rc, lane);
  else eval_reactions<W>(N, p, L, lane, true);
  const unsigned *sq = N.jsched + lane;
  double acc = 0.0;
#pragma unroll 4
  for (int s = 0; s < N.jslots; ++s) {
    const unsigned w = sq[64 * s];
    const double v = L.part[w & 0xFFu];
    const double c = (double)(int)(int8_t)((w >> 8) & 0xFFu);
    const double nacc = acc + c * v;
    acc = (w >> 26) & 1u ? nacc : acc;
    const bool last = (w >> 25) & 1u;
    L.J[last ? (int)((w >> 16) & 0x1FFu) : wave::kNNZ] = acc;
    acc = last ? 0.0 : acc;
  }
  __syncthreads();
}

// dy/dt, balanced over the lanes: the entries are cut into 128 shares; lane
// l sums shares l, l + W, ... piece by piece (each piece in the upstream
// order) into psum, then species s adds its pieces in order.  A species
// whose row lies in one piece gets the bits of rhs_lane; a row split over
// shares differs by the rounding of the split.  Requires y published.
template <int W>
__device__ double rhs_balanced(const NetDev &N, const CellParams &p, const Lds &L, int lane) {
  eval_reactions<W>(N, p, L, lane, false);
  for (int sh = lane; sh < kMaxSpeciesWide; sh += W) {
    for (int q = N.pc_lane[sh]; q < N.pc_lane[sh + 1]; ++q) {
      const int e1 = N.pc_beg[q + 1];
      int e = N.pc_beg[q];
      double acc = 0.0;
      for (; e + 4 <= e1; e += 4) {
        const int c0 = N.rhs_pk[e], c1 = N.rhs_pk[e + 1], c2 = N.rhs_pk[e + 2],
                  c3 = N.rhs_pk[e + 3];
        const double v0 = L.part[2 * (c0 >> 1)], v1 = L.part[2 * (c1 >> 1)],
                     v2 = L.part[2 * (c2 >> 1)], v3 = L.part[2 * (c3 >> 1)];
        acc = (c0 & 1) ? acc + v0 : acc - v0;
        acc = (c1 & 1) ? acc + v1 : acc - v1;
        acc = (c2 & 1) ? acc + v2 : acc - v2;
        acc = (c3 & 1) ? acc + v3 : acc - v3;
      }
      for (; e < e1; ++e) {
        const int c0 = N.rhs_pk[e];
        const double v0 = L.part[2 * (c0 >> 1)];
        acc = (c0 & 1) ? acc + v0 : acc - v0;
      }
      L.psum[q] = acc;
    }
  }
  __syncthreads();
  double acc = 0.0;
  if (lane < N.ns)
    for (int q = N.pc_row[lane]; q < N.pc_row[lane + 1]; ++q) acc += L.psum[q];
  __syncthreads();  // part[] and psum[] are reused by the next evaluation
  return acc;
}

template <class G>
__device__ __forceinline__ void jac_eval(const NetDev &N, const CellParams &p, const Lds &L,
                                         int lane) {
  if constexpr (G::W == 64) jacobian(N, p, L, lane);
  else jacobian_sparse<G::W>(N, p, L, lane);
}

__device__ void lu_factor_inplace(const NetDev &N, const Lds &L, int lane);

// M = I - c J, factorised in place with partial pivoting.
__device__ void lu_factor(const NetDev &N, const Lds &L, double c, int lane) {
  const int n = N.ns, ld = L.ld;
  for (int idx = lane; idx < n * ld; idx += 64) {
    int i = idx / ld, j = idx - i * ld;
    L.M[idx] = (i == j ? 1.0 : 0.0) - c * L.J[idx];
  }
  __syncthreads();
  lu_factor_inplace(N, L, lane);
}

// L.M factorised in place with partial pivoting (row interchanges in L.piv).
__device__ void lu_factor_inplace(const NetDev &N, const Lds &L, int lane) {
  const int n = N.ns, ld = L.ld;
  const bool half = n <= 32;
  const int col = half ? (lane & 31) : lane;
  const int r0 = half ? (lane >> 5) : 0, rs = half ? 2 : 1;
  for (int k = 0; k < n; ++k) {
    double v = (lane >= k && lane < n) ? fabs(L.M[lane * ld + k]) : -1.0;
    const int pk = wave_argmax(v, lane);
    if (pk != k && lane < n) {
      double a = L.M[k * ld + lane];
      L.M[k * ld + lane] = L.M[pk * ld + lane];
      L.M[pk * ld + lane] = a;
    }
    if (lane == 0) L.piv[k] = pk;
    __syncthreads();
    const double inv = 1.0 / L.M[k * ld + k];
    if (lane > k && lane < n) L.M[lane * ld + k] *= inv;
    __syncthreads();
    if (col > k && col < n) {
      const double ukj = L.M[k * ld + col];
      for (int i = k + 1 + r0; i < n; i += rs) L.M[i * ld + col] -= L.M[i * ld + k] * ukj;
    }
    __syncthreads();
  }
}

// Solve M x = b with the factors (b and x live in registers, lane = row).
// The pivots and the lane's row entries are read four per LDS round trip,
// ahead of the readlane chain of the substitutions.
constexpr int kSolveChunk = 4;
__device__ double lu_solve(const NetDev &N, const Lds &L, double b, int lane) {
  const int n = N.ns, ld = L.ld;
  const double *Mr = L.M + min(lane, n - 1) * ld;  // rows past n: entries unused
  for (int k0 = 0; k0 < n; k0 += kSolveChunk) {
    int pv[kSolveChunk];
#pragma unroll
    for (int q = 0; q < kSolveChunk; ++q) pv[q] = L.piv[min(k0 + q, n - 1)];
#pragma unroll
    for (int q = 0; q < kSolveChunk; ++q) {
      const int k = k0 + q, pk = uniform(pv[q]);
      if (k < n && pk != k) {
        double bk = bcast(b, k), bp = bcast(b, pk);
        if (lane == k) b = bp;
        if (lane == pk) b = bk;
      }
    }
  }
  for (int k0 = 0; k0 < n - 1; k0 += kSolveChunk) {
    double a[kSolveChunk];
#pragma unroll
    for (int q = 0; q < kSolveChunk; ++q) a[q] = Mr[min(k0 + q, n - 1)];
#pragma unroll
    for (int q = 0; q < kSolveChunk; ++q) {
      const int k = k0 + q;
      if (k < n - 1) {
        double bk = bcast(b, k);
        if (lane > k && lane < n) b -= a[q] * bk;
      }
    }
  }
  for (int k0 = n - 1; k0 >= 0; k0 -= kSolveChunk) {
    double a[kSolveChunk];
#pragma unroll
    for (int q = 0; q < kSolveChunk; ++q) a[q] = Mr[max(k0 - q, 0)];
#pragma unroll
    for (int q = 0; q < kSolveChunk; ++q) {
      const int k = k0 - q;
      if (k >= 0) {
        if (lane == k) b /= a[q];  // the lane's own diagonal entry
        double bk = bcast(b, k);
        if (lane < k) b -= a[q] * bk;
      }
    }
  }
  return b;
}

// lu_solve for the two-wavefront form, run by wavefront 0 alone: lane l
// holds rows l and l + 64 and the solved components travel by readlane, as
// in lu_solve (same operations in the same order, the same chunked reads);
// b enters and leaves through an LDS vector, two barriers per solve.
__device__ double lu_solve_wide(const NetDev &N, const Lds &L, double b, int lane) {
  const int n = N.ns, ld = L.ld;
  double *x = L.xs;
  if (lane < n) x[lane] = b;
  __syncthreads();
  if (lane < 64) {
    const int r0 = lane, r1 = lane + 64;
    const double *M0 = L.M + min(r0, n - 1) * ld, *M1 = L.M + min(r1, n - 1) * ld;
    double b0 = r0 < n ? x[r0] : 0.0, b1 = r1 < n ? x[r1] : 0.0;
    for (int k0 = 0; k0 < n; k0 += kSolveChunk) {
      int pv[kSolveChunk];
#pragma unroll
      for (int q = 0; q < kSolveChunk; ++q) pv[q] = L.piv[min(k0 + q, n - 1)];
#pragma unroll
      for (int q = 0; q < kSolveChunk; ++q) {
        const int k = k0 + q, pk = uniform(pv[q]);
        if (k < n && pk != k) {
          const double bk = bcast(k < 64 ? b0 : b1, k & 63);
          const double bp = bcast(pk < 64 ? b0 : b1, pk & 63);
          if (lane == (k & 63)) {
            if (k < 64) b0 = bp;
            else b1 = bp;
          }
          if (lane == (pk & 63)) {
            if (pk < 64) b0 = bk;
            else b1 = bk;
          }
        }
      }
    }
    for (int k0 = 0; k0 < n - 1; k0 += kSolveChunk) {
      double a0[kSolveChunk], a1[kSolveChunk];
#pragma unroll
      for (int q = 0; q < kSolveChunk; ++q) {
        const int k = min(k0 + q, n - 1);
        a0[q] = M0[k];
        a1[q] = M1[k];
      }
#pragma unroll
      for (int q = 0; q < kSolveChunk; ++q) {
        const int k = k0 + q;
        if (k < n - 1) {
          const double bk = bcast(k < 64 ? b0 : b1, k & 63);
          if (r0 > k && r0 < n) b0 -= a0[q] * bk;
          if (r1 > k && r1 < n) b1 -= a1[q] * bk;
        }
      }
    }
    for (int k0 = n - 1; k0 >= 0; k0 -= kSolveChunk) {
      double a0[kSolveChunk], a1[kSolveChunk];
#pragma unroll
      for (int q = 0; q < kSolveChunk; ++q) {
        const int k = max(k0 - q, 0);
        a0[q] = M0[k];
        a1[q] = M1[k];
      }
#pragma unroll
      for (int q = 0; q < kSolveChunk; ++q) {
        const int k = k0 - q;
        if (k >= 0) {
          if (lane == (k & 63)) {  // the lane's own diagonal entry
            if (k < 64) b0 /= a0[q];
            else b1 /= a1[q];
          }
          const double bk = bcast(k < 64 ? b0 : b1, k & 63);
          if (r0 < k) b0 -= a0[q] * bk;
          if (r1 < k) b1 -= a1[q] * bk;
        }
      }
    }
    x[r0] = b0;
    x[r1] = b1;
  }
  __syncthreads();
  b = lane < n ? x[lane] : 0.0;
  __syncthreads();  // x is reused by the next solve
  return b;
}

// Linear-algebra policies of the Newton iteration matrix M = I - c J.
//  LinLds: J and M in LDS, any n <= 64 (lane = column in the elimination).
//  LinReg: n <= 32; lane i keeps row i of J and of M in registers.  The
//          elimination never moves rows between lanes: step k picks the
//          pivot lane p_k by a wavefront argmax, reads the pivot row with
//          readlane (scalar operands for every lane) and updates the rows not
//          yet used as pivots with plain FMAs.  The solves replay the same
//          lane order and return component j on lane j.
struct LinLds {
  __device__ void load_jac(const NetDev &, const Lds &, int) {}
  __device__ void factor(const NetDev &N, const Lds &L, double c, int lane) {
    lu_factor(N, L, c, lane);
  }
  __device__ double solve(const NetDev &N, const Lds &L, double b, int lane) {
    return lu_solve(N, L, b, lane);
  }
};

// LinMfma: M = I - c J in LDS (n <= 64), factorised with partial pivoting
// in panels of four columns (LAPACK's blocked right-looking order: panel
// with row interchanges across all columns, U12 = L11^-1 A12, then
// A22 -= L21 U12).  The trailing update is a rank-4 product per 16 x 16
// tile, one v_mfma_f64_16x16x4f64 each.  Operand layout of that
// instruction (64 lanes): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator element r of lane l is D[4r + l / 16][l % 16].
// The factors have the layout of lu_factor's (unit L below the diagonal, U
// on and above it, interchanges in L.piv), so lu_solve applies.
// G: the workgroup (one wavefront, or two for up to 128 species, which
// share the trailing tiles by checkerboard).  SparseJ: J holds the
// structural entries only (two-wavefront LDS layout).
typedef double p3d_f64x4 __attribute__((ext_vector_type(4)));

template <class G, bool SparseJ>
struct LinMfmaG {
  __device__ void load_jac(const NetDev &, const Lds &, int) {}

  __device__ void factor(const NetDev &N, const Lds &L, double c, int lane) {
    const int n = N.ns, ld = L.ld;
    const G g{L.red};
    if constexpr (SparseJ) {
      for (int idx = lane; idx < n * ld; idx += G::W) {
        const int i = idx / ld, j = idx - i * ld;
        L.M[idx] = i == j ? 1.0 : 0.0;
      }
      __syncthreads();
      for (int z = lane; z < N.nnz; z += G::W) {
        const int i = N.jrow[z], j = N.jcol[z];
        L.M[i * ld + j] = (i == j ? 1.0 : 0.0) - c * L.J[z];
      }
    } else {
      for (int idx = lane; idx < n * ld; idx += G::W) {
        const int i = idx / ld, j = idx - i * ld;
        L.M[idx] = (i == j ? 1.0 : 0.0) - c * L.J[idx];
      }
    }
    __syncthreads();
    const int nt = (n + 15) / 16;
    const int lr = lane & 15, lq = (lane & 63) >> 4, wv = lane >> 6;
    for (int k0 = 0; k0 < n; k0 += 4) {
      const int k1 = min(k0 + 4, n);
      // panel columns k0..k1-1, unblocked; lane = row for the column work,
      // lane = column for the interchanges
      for (int k = k0; k < k1; ++k) {
        const double v = (lane >= k && lane < n) ? fabs(L.M[lane * ld + k]) : -1.0;
        const int pk = g.argmax(v, lane);
        if (pk != k && lane < n) {
          const double a = L.M[k * ld + lane];
          L.M[k * ld + lane] = L.M[pk * ld + lane];
          L.M[pk * ld + lane] = a;
        }
        if (lane == 0) L.piv[k] = pk;
        __syncthreads();
        const double inv = 1.0 / L.M[k * ld + k];
        if (lane > k && lane < n) {
          const double lik = L.M[lane * ld + k] * inv;
          L.M[lane * ld + k] = lik;
          for (int j = k + 1; j < k1; ++j) L.M[lane * ld + j] -= lik * L.M[k * ld + j];
        }
        __syncthreads();
      }
      if (k1 >= n) break;
      // U12 = L11^-1 A12 (lane = column)
      if (lane >= k1 && lane < n) {
        for (int r = k0 + 1; r < k1; ++r) {
          double acc = L.M[r * ld + lane];
          for (int q = k0; q < r; ++q) acc -= L.M[r * ld + q] * L.M[q * ld + lane];
          L.M[r * ld + lane] = acc;
        }
      }
      __syncthreads();
      // A22 -= L21 U12, tile by tile; operands outside rows/columns
      // k1..n-1 or beyond the panel are zero, so the tiles need no alignment
      const int q = k0 + lq;
      for (int ti = k1 / 16; ti < nt; ++ti) {
        const int ia = ti * 16 + lr;
        const double a = (ia >= k1 && ia < n && q < k1) ? -L.M[ia * ld + q] : 0.0;
        for (int tj = k1 / 16; tj < nt; ++tj) {
          if (G::W > 64 && ((ti + tj) & 1) != wv) continue;
          const int jb = tj * 16 + lr;
          const double b = (jb >= k1 && jb < n && q < k1) ? L.M[q * ld + jb] : 0.0;
          p3d_f64x4 acc;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = ti * 16 + 4 * r + lq, col = tj * 16 + lr;
            acc[r] = (row < n && col < n) ? L.M[row * ld + col] : 0.0;
          }
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = ti * 16 + 4 * r + lq, col = tj * 16 + lr;
            if (row >= k1 && row < n && col >= k1 && col < n) L.M[row * ld + col] = acc[r];
          }
        }
      }
      __syncthreads();
    }
  }

  __device__ double solve(const NetDev &N, const Lds &L, double b, int lane) {
    if constexpr (G::W == 64) return lu_solve(N, L, b, lane);
    else return lu_solve_wide(N, L, b, lane);
  }
};
using LinMfma = LinMfmaG<Wave1, false>;

// LinSparseG: M = I - c J in LDS in the network's static elimination order
// (minimum Markowitz product, no pivoting).  Lane = species s, whose row
// sits at position p = sl_pos[s].  The schedule is staged in LDS once per
// cell (stage()).  The factorisation is right-looking over dependency
// levels: in pass v every wavefront walks the pivots of level v in position
// order (each pivot's descriptor made wave-uniform by readfirstlane, its
// upper-row positions read as bytes by broadcast LDS reads), and each lane
// whose row has a multiplier for the pivot applies it (a row without that
// entry holds an exact zero there, kept from the zeroed row, and skips it).  A pivot's level is above those of
// every pivot sharing an entry of its row or column, so a pass reads only
// rows that earlier passes finished; a row inverts its diagonal at the end
// of the pass before its own level.  One barrier per level.  Factors: unit
// L below the diagonal, U above it, 1 / u_pp on it, zeros outside the
// filled structure, so the substitutions are the dense serial ones in
// position order, run by wavefront 0 as in lu_solve_wide.
template <class G>
struct LinSparseG {
  __device__ void load_jac(const NetDev &, const Lds &, int) {}

  // the schedule into LDS (before the first factorisation; every lane calls)
  __device__ static void stage(const NetDev &N, const Lds &L, int lane) {
    for (int i = lane; i < N.ns; i += G::W) L.sd[i] = N.sl_desc[i];
    unsigned *su = reinterpret_cast<unsigned *>(L.su);
    for (int i = lane; i < (N.sl_nub >> 2); i += G::W) su[i] = N.sl_ub[i];
    unsigned *sz = reinterpret_cast<unsigned *>(L.sz);
    for (int i = lane; i < (N.sl_nzb >> 2); i += G::W) sz[i] = N.sl_zb[i];
    unsigned *sp = reinterpret_cast<unsigned *>(L.sp);
    for (int i = lane; i < (kMaxSpeciesWide >> 2); i += G::W) sp[i] = N.sl_pb[i];
  }

  __device__ void factor(const NetDev &N, const Lds &L, double c, int lane) {
    const int ld = L.ld, ns = N.ns;
    const bool act = lane < ns;
    const int p = act ? (int)L.sp[lane] : ns;  // lanes past the species: no row
    const int lev = act ? N.sl_lev[p] : -2;
    double *Mp = L.M + (act ? p : 0) * ld;
    if (act) {
      const int z0 = N.jr_start[lane], z1 = N.jr_start[lane + 1];
      for (int j = 0; j < ns; ++j) Mp[j] = 0.0;
      Mp[p] = 1.0;
      for (int z = z0; z < z1; ++z) {
        const int j = L.sz[z];
        Mp[j] = (j == p ? 1.0 : 0.0) - c * L.J[z];
      }
      if (lev == 0) Mp[p] = 1.0 / Mp[p];
    }
    __syncthreads();
    int v = 0;
    for (int q = 0; q < ns; ++q) {
      // wave-uniform descriptor: scalar fields and branches
      const unsigned d = __builtin_amdgcn_readfirstlane(L.sd[q]);
      const int k = (int)(d & 127u), len = (int)((d >> 7) & 127u);
      const int ub = (int)((d >> 14) & 0x1ffffu);
      const double *Mk = L.M + k * ld;
      double l = 0.0;
      if (p > k && p < ns) l = Mp[k];
      if (l != 0.0) {
        l *= Mk[k];
        Mp[k] = l;
        const unsigned *Uw = reinterpret_cast<const unsigned *>(L.su + ub);
        int w = 0;
        for (; w + 4 <= len; w += 4) {
          const unsigned jj = Uw[w >> 2];
          const int j0 = jj & 0xff, j1 = (jj >> 8) & 0xff, j2 = (jj >> 16) & 0xff, j3 = jj >> 24;
          const double b0 = Mk[j0], b1 = Mk[j1], b2 = Mk[j2], b3 = Mk[j3];
          const double a0 = Mp[j0], a1 = Mp[j1], a2 = Mp[j2], a3 = Mp[j3];
          Mp[j0] = a0 - l * b0;
          Mp[j1] = a1 - l * b1;
          Mp[j2] = a2 - l * b2;
          Mp[j3] = a3 - l * b3;
        }
        if (w < len) {
          const unsigned jj = Uw[w >> 2];
          for (int t = 0; t < len - w; ++t) {
            const int j = (jj >> (8 * t)) & 0xff;
            Mp[j] -= l * Mk[j];
          }
        }
      }
      if (d >> 31) {  // the last pivot of level v
        if (lev == v + 1) Mp[p] = 1.0 / Mp[p];
        __syncthreads();
        ++v;
      }
    }
  }

  __device__ double solve(const NetDev &N, const Lds &L, double b, int lane) {
    const int n = N.ns, ld = L.ld;
    double *x = L.xs;
    if (lane < n) x[L.sp[lane]] = b;
    __syncthreads();
    if (lane < 64) {
      const int r0 = lane, r1 = lane + 64;
      const double *M0 = L.M + min(r0, n - 1) * ld, *M1 = L.M + min(r1, n - 1) * ld;
      double b0 = r0 < n ? x[r0] : 0.0, b1 = r1 < n ? x[r1] : 0.0;
      for (int k0 = 0; k0 < n - 1; k0 += kSolveChunk) {
        double a0[kSolveChunk], a1[kSolveChunk];
#pragma unroll
        for (int q = 0; q < kSolveChunk; ++q) {
          const int k = min(k0 + q, n - 1);
          a0[q] = M0[k];
          a1[q] = M1[k];
        }
#pragma unroll
        for (int q = 0; q < kSolveChunk; ++q) {
          const int k = k0 + q;
          if (k < n - 1) {
            const double bk = bcast(k < 64 ? b0 : b1, k & 63);
            if (r0 > k && r0 < n) b0 -= a0[q] * bk;
            if (r1 > k && r1 < n) b1 -= a1[q] * bk;
          }
        }
      }
      for (int k0 = n - 1; k0 >= 0; k0 -= kSolveChunk) {
        double a0[kSolveChunk], a1[kSolveChunk];
#pragma unroll
        for (int q = 0; q < kSolveChunk; ++q) {
          const int k = max(k0 - q, 0);
          a0[q] = M0[k];
          a1[q] = M1[k];
        }
#pragma unroll
        for (int q = 0; q < kSolveChunk; ++q) {
          const int k = k0 - q;
          if (k >= 0) {
            if (lane == (k & 63)) {  // the lane's own reciprocal diagonal
              if (k < 64) b0 *= a0[q];
              else b1 *= a1[q];
            }
            const double bk = bcast(k < 64 ? b0 : b1, k & 63);
            if (r0 < k) b0 -= a0[q] * bk;
            if (r1 < k) b1 -= a1[q] * bk;
          }
        }
      }
      x[r0] = b0;
      if (r1 < kMaxSpeciesWide) x[r1] = b1;
    }
    __syncthreads();
    b = lane < n ? x[L.sp[lane]] : 0.0;
    __syncthreads();  // x is reused by the next solve
    return b;
  }
};

// LinLdsU: M in LDS for n <= NP (32), every loop fully unrolled so that the
// LDS loads of a step issue together, ahead of the dependency chain.  Lane L
// handles column L & 31 of the trailing update on rows of parity L >> 5.
template <int NP>
struct LinLdsU {
  __device__ void load_jac(const NetDev &, const Lds &, int) {}

  __device__ __forceinline__ void factor(const NetDev &N, const Lds &L, double c, int lane) {
    const int n = N.ns, ld = L.ld;
    for (int idx = lane; idx < n * ld; idx += 64) {
      const int i = idx / ld, j = idx - i * ld;
      L.M[idx] = (i == j ? 1.0 : 0.0) - c * L.J[idx];
    }
    __syncthreads();
    const int col = lane & 31, par = lane >> 5;
#pragma clang loop unroll(full)
    for (int k = 0; k < NP; ++k) {
      if (k < n) {
        const double v = (lane >= k && lane < n) ? fabs(L.M[lane * ld + k]) : -1.0;
        const int pk = wave_argmax(v, lane);
        if (pk != k && lane < n) {
          const double a = L.M[k * ld + lane];
          L.M[k * ld + lane] = L.M[pk * ld + lane];
          L.M[pk * ld + lane] = a;
        }
        if (lane == 0) L.piv[k] = pk;
        __syncthreads();
        const double inv = 1.0 / L.M[k * ld + k];
        const double ukj = (col < n) ? L.M[k * ld + col] : 0.0;
        double lik[NP];
#pragma clang loop unroll(full)
        for (int i = k + 1; i < NP; ++i) lik[i] = (i < n) ? L.M[i * ld + k] * inv : 0.0;
#pragma clang loop unroll(full)
        for (int i = k + 1; i < NP; ++i) {
          if (i < n && ((i - k - 1) & 1) == par) {
            if (col == k) L.M[i * ld + k] = lik[i];
            else if (col > k && col < n) L.M[i * ld + col] -= lik[i] * ukj;
          }
        }
        __syncthreads();
      }
    }
  }

  __device__ __forceinline__ double solve(const NetDev &N, const Lds &L, double b, int lane) {
    const int n = N.ns, ld = L.ld;
    double mrow[NP];
#pragma clang loop unroll(full)
    for (int k = 0; k < NP; ++k) mrow[k] = (lane < n && k < n) ? L.M[lane * ld + k] : 0.0;
#pragma clang loop unroll(full)
    for (int k = 0; k < NP; ++k) {
      if (k < n) {
        const int pk = uniform(L.piv[k]);
        if (pk != k) {
          const double bk = bcast(b, k), bp = bcast(b, pk);
          if (lane == k) b = bp;
          if (lane == pk) b = bk;
        }
      }
    }
#pragma clang loop unroll(full)
    for (int k = 0; k < NP - 1; ++k) {
      if (k < n - 1) {
        const double bk = bcast(b, k);
        if (lane > k && lane < n) b -= mrow[k] * bk;
      }
    }
#pragma clang loop unroll(full)
    for (int k = NP - 1; k >= 0; --k) {
      if (k < n) {
        if (lane == k) b /= mrow[k];
        const double bk = bcast(b, k);
        if (lane < k) b -= mrow[k] * bk;
      }
    }
    return b;
  }
};

template <int NP>
struct LinReg {
  // rows >= n are identity rows, so the elimination always runs over NP;
  // the pivot lane of step k is kept in LDS (L.piv[k])
  double Mr[NP];
  int step_of;  // elimination step at which this lane's row became the pivot
  int src;      // lane holding solution component `lane`

  __device__ void load_jac(const NetDev &, const Lds &, int) {}

  __device__ __forceinline__ void factor(const NetDev &N, const Lds &L, double c, int lane) {
    const int n = N.ns;
    const bool row = lane < n;
    const double *Jrow = L.J + (row ? lane : 0) * L.ld;
#pragma clang loop unroll(full)
    for (int j = 0; j < NP; ++j) {
      const double jv = (row && j < n) ? Jrow[j] : 0.0;
      Mr[j] = (j == lane ? 1.0 : 0.0) - c * jv;
    }
    step_of = NP;
    src = lane;
#pragma clang loop unroll(full)
    for (int k = 0; k < NP; ++k) {
      const double v = (lane < NP && step_of == NP) ? fabs(Mr[k]) : -1.0;
      const int p = wave_argmax(v, lane);
      if (lane == 0) L.piv[k] = p;
      if (lane == p) step_of = k;
      if (lane == k) src = p;
      const double inv = 1.0 / bcast(Mr[k], p);
      const bool upd = step_of == NP;
      const double l = Mr[k] * inv;
      if (upd) Mr[k] = l;
#pragma clang loop unroll(full)
      for (int j = k + 1; j < NP; ++j) {
        const double u = bcast(Mr[j], p);
        if (upd) Mr[j] -= l * u;
      }
    }
    __syncthreads();
  }

  __device__ __forceinline__ double solve(const NetDev &N, const Lds &L, double b, int lane) {
    if (lane >= N.ns) b = 0.0;
#pragma clang loop unroll(full)
    for (int k = 0; k < NP; ++k) {  // forward: unit-lower factor in pivot order
      const double yk = bcast(b, uniform(L.piv[k]));
      if (step_of > k) b -= Mr[k] * yk;
    }
#pragma clang loop unroll(full)
    for (int k = NP - 1; k >= 0; --k) {  // backward: upper factor
      const int pk = uniform(L.piv[k]);
      const double q = b / Mr[k];
      if (lane == pk) b = q;
      const double xk = bcast(b, pk);
      if (step_of < k) b -= Mr[k] * xk;
    }
    return __shfl(b, src, 64);
  }
};

// The bundled network's no-pivot LU in registers (wave_net.h): lane =
// species holds its row of M = I - c J by elimination position; the
// factorisation broadcasts only the structural entries of each pivot row
// (readlane), in the minimum-Markowitz order of the one-cell-per-lane code,
// so no pivot search and no LDS traffic.  Needs the network structure the
// code was generated for (Network::lane_ok).
__constant__ unsigned c_wave_lmask[64] = P3D_WAVE_LMASK;
__constant__ unsigned c_wave_umask[64] = P3D_WAVE_UMASK;
__constant__ int c_wave_pos[64] = P3D_WAVE_POS;
__constant__ int c_wave_lev[64] = P3D_WAVE_LEVELS;

struct WaveBcast {
  __device__ __forceinline__ double operator()(double v, int src) const { return bcast(v, src); }
};

// The lane's row of M in LDS as the generated factorisation's `a`: a[j]
// reads or writes entry j of the row, and the broadcast of a[j] from the
// pivot lane reads entry j of the pivot's row.  The statements round as the
// register form (a[j] = a[j] - m * u is one FMA either way).
struct LdsEntry {
  double *row;
  int j;
  __device__ __forceinline__ operator double() const { return row[j]; }
  __device__ __forceinline__ LdsEntry &operator=(double v) {
    row[j] = v;
    return *this;
  }
};
struct LdsRow {
  double *row;
  __device__ __forceinline__ LdsEntry operator[](int j) const { return LdsEntry{row, j}; }
};
struct LdsRowBcast {
  const double *M;
  int ld;
  __device__ __forceinline__ double operator()(const LdsEntry &e, int src) const {
    return M[src * ld + e.j];
  }
};

struct LinStatic {
  // M and its factors live in LDS (the lane's row, by position, in L.M): the
  // factorisation updates the rows in place (no 32-entry register row, which
  // halved the kernel's scratch spills and measured 4.6% faster), and each
  // solve loads the lane's row of the factors.
  static constexpr bool kNeedsM = true;
  unsigned pw0, pw1;  // this lane's RHS piece (rhs_pieces)
  int pos, lev;       // this lane's elimination position, solve and factorisation levels
  __device__ LinStatic() {
    pw0 = c_wave_pw0[threadIdx.x & 63];
    pw1 = c_wave_pw1[threadIdx.x & 63];
    pos = c_wave_pos[threadIdx.x & 63];
    lev = c_wave_lev[threadIdx.x & 63];
  }

  __device__ void load_jac(const NetDev &, const Lds &, int) {}

  // J is sparse in LDS (carve_wave): the lane's row entries start at
  // N.jr_start[lane], one per set bit of N.jr_mask[lane] in column order.
  __device__ __forceinline__ void factor(const NetDev &N, const Lds &L, double c, int lane) {
    const bool row = lane < wave::kNS;
    const unsigned lm = row ? c_wave_lmask[lane] : 0u;
    const unsigned jm = row ? N.jr_mask[lane] : 0u;
    const int js = row ? N.jr_start[lane] : 0;
    // lanes past the species write a spare row (kNS), so the loop has no branches
    double *Mrow = L.M + (row ? lane : wave::kNS) * L.ld;
#pragma unroll
    for (int q = 0; q < wave::kNS; ++q) {
      const int col = wave::kPerm[q];
      const bool has = (jm >> col) & 1u;
      const int z = has ? js + __builtin_popcount(jm & ((1u << col) - 1u)) : js;
      const double jz = L.J[z];
      const double jv = has ? jz : 0.0;
      Mrow[q] = (col == lane ? 1.0 : 0.0) - c * jv;
    }
    LdsRow ar{Mrow};
    wave::lu_factor(ar, lm, lane, LdsRowBcast{L.M, L.ld}, pos, lev);
  }

  // The solves take the lane's row of the factors into registers first: the
  // substitution chain is latency-bound, and reading each entry from LDS as
  // it is used measured 17% slower.
  __device__ __forceinline__ double solve(const NetDev &N, const Lds &L, double b, int lane) {
    const bool row = lane < wave::kNS;
    const unsigned lm = row ? c_wave_lmask[lane] : 0u;
    const unsigned um = row ? c_wave_umask[lane] : 0u;
    // lanes past the species read the spare row: their values are never used
    const double *Mrow = L.M + (row ? lane : wave::kNS) * L.ld;
    double a[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) a[q] = q < wave::kNS ? Mrow[q] : 0.0;
    const double dinv = Mrow[pos < wave::kNS ? pos : 0];  // the factor's 1 / u_ii
    if (lane >= N.ns) b = 0.0;
    const double x = wave::lu_solve(a, lm, um, lane, b, WaveBcast{}, L.xs, pos, lev, dinv);
    return row ? x : 0.0;
  }
};

__device__ __forceinline__ double rms(double v, int ns) { return sqrt(wave_sum(v * v) / ns); }

// The Jacobian in the layout the linear-algebra form reads: sparse for the
// register LU (LinStatic, carve_wave), otherwise jac_eval's.
template <class Lin, class G>
__device__ __forceinline__ void jac_eval_for(const NetDev &N, const CellParams &p, const Lds &L,
                                             const RxCache &rc, int lane) {
  if constexpr (std::is_same<Lin, LinStatic>::value) jacobian_sched<64>(N, p, L, lane, &rc);
  else jac_eval<G>(N, p, L, lane);
}

// Step-size change of the backward differences D[0..order] by `factor`:
// D <- (R(factor) U)^T D with R_im = prod_{q=1..i} (q-1 - factor m)/q and
// U = R(1).  Lanes 0..35 build RU (entry (i, j) per lane) in LDS, then every
// species lane applies it to its own differences.
// U = R(1) of change_D: U_mj = prod_{q=1..m} (q-1 - j)/q, formed at compile
// time with the same operations in the same order as the loop it replaces
constexpr double bdf_u(int m, int j) {
  double u = 1.0;
  for (int q = 1; q <= m; ++q) u *= ((double)(q - 1) - (double)j) / (double)q;
  return u;
}
#define P3D_U6(m) bdf_u(m, 0), bdf_u(m, 1), bdf_u(m, 2), bdf_u(m, 3), bdf_u(m, 4), bdf_u(m, 5)
__constant__ double c_bdf_u[36] = {P3D_U6(0), P3D_U6(1), P3D_U6(2), P3D_U6(3), P3D_U6(4),
                                   P3D_U6(5)};
#undef P3D_U6

__device__ void change_D(const Lds &L, int lane, double (&D)[kNumD], int order, double factor) {
  // R_im once per lane (i = l / 6, m = l % 6 of the wavefront's lane l: at
  // most 5 divisions), then lane (i, j) gathers R_i0..R_i5 from lanes 6i..6i+5
  // and forms (RU)_ij = sum_m R_im U_mj in the same order as before
  const int wl = lane & 63;
  double r = 1.0;
  {
    const int i = wl / 6, m = wl % 6;
    for (int q = 1; q <= i && q <= kMaxOrder; ++q)
      r *= ((double)(q - 1) - factor * (double)m) / (double)q;
  }
  double rim[kMaxOrder + 1];
#pragma unroll
  for (int m = 0; m <= kMaxOrder; ++m) rim[m] = __shfl(r, 6 * (wl / 6) + m, 64);
  if (lane < 36) {
    const int i = lane / 6, j = lane % 6;
    double acc = 0.0;
    if (i <= order && j <= order) {
#pragma unroll
      for (int m = 0; m <= kMaxOrder; ++m)
        if (m <= order) acc += rim[m] * c_bdf_u[6 * m + j];
    }
    L.ru[lane] = acc;
  }
  __syncthreads();
  double Dn[kMaxOrder + 1];
#pragma unroll
  for (int j = 0; j <= kMaxOrder; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i <= kMaxOrder; ++i) acc += L.ru[6 * i + j] * D[i];
    Dn[j] = acc;
    // keep the 36 RU loads from being hoisted into registers all at once
    __asm__ volatile("" ::: "memory");
  }
#pragma unroll
  for (int j = 0; j <= kMaxOrder; ++j)
    if (j <= order) D[j] = Dn[j];
  __syncthreads();
}

// NDF coefficients (Shampine & Reichelt 1997): kappa_k, gamma_k = sum 1/j,
// alpha_k = (1 - kappa_k) gamma_k, error constants kappa_k gamma_k + 1/(k+1).
__constant__ double c_gam[kMaxOrder + 2] = {0.0, 1.0, 1.5, 1.8333333333333333,
                                            2.0833333333333333, 2.2833333333333333,
                                            2.45};
__constant__ double c_kappa[kMaxOrder + 1] = {0.0, -0.1850, -1.0 / 9.0, -0.0823, -0.0415, 0.0};

struct BdfStats {
  int steps, rhs, jac, lu, rejects;
};

// Step recorder of bdf_integrate.  NoRec records nothing (every other form);
// the taped forward of the discrete adjoint records the step sequence.
struct NoRec {
  __device__ __forceinline__ void init(const double (&)[kNumD], double, int) {}
  __device__ __forceinline__ void change(double) {}
  __device__ __forceinline__ void accept(int, double, const double (&)[kNumD], int) {}
};

// Variable-order BDF from t0 to t1 for the wave's cell.  y: lane's species.
// Returns 0 on success, -1 step underflow, -2 step limit.
template <class Lin, class G, class Rec>
__device__ __forceinline__ int bdf_integrate_rec(const NetDev &N, const CellParams &p,
                                                 const Lds &L, Lin &lin, const RxCache &rc,
                                                 int lane, double &y, double t0, double t1,
                                                 double rtol, double atol, double max_step,
                                                 int max_steps, BdfStats &st, Rec &rec) {
  const G g{L.red};
  const bool cached = G::W == 64 && N.nr <= 64 * kRxPerLane;
  auto rhs = [&]() {
    if constexpr (std::is_same<Lin, LinStatic>::value)
      return rhs_pieces(N, p, L, rc, lin.pw0, lin.pw1, lane);
    else if constexpr (G::W == 64)
      return cached ? rhs_cached(N, p, L, rc, lane) : rhs_balanced<64>(N, p, L, lane);
    else
      return rhs_balanced<G::W>(N, p, L, lane);
  };
  const int ns = N.ns;
  const bool act = lane < ns;
  // coefficients indexed by the (wave-uniform) order: scalar loads
  auto gam = [](int k) { return c_gam[k]; };
  auto alp = [](int k) { return (1.0 - c_kappa[k]) * c_gam[k]; };
  auto ec = [](int k) { return c_kappa[k] * c_gam[k] + 1.0 / (k + 1); };

  const double span = t1 - t0;
  if (!(span > 0.0)) return 0;
  if (max_step <= 0.0 || max_step > span) max_step = span;
  if (!act) y = 0.0;

  publish_y(N, L, y, lane);
  double f0 = rhs();
  st.rhs++;
  // initial step (Hairer, Norsett & Wanner 1993, II.4), order 1
  double scale = atol + rtol * fabs(y);
  double d0 = rms_g(g, act ? y / scale : 0.0, ns), d1 = rms_g(g, act ? f0 / scale : 0.0, ns);
  double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
  h0 = fmin(h0, span);
  publish_y(N, L, y + h0 * f0, lane);
  double f1 = rhs();
  st.rhs++;
  double d2 = rms_g(g, act ? (f1 - f0) / scale : 0.0, ns) / h0;
  double dm = fmax(d1, d2);
  double h1 = (dm <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : sqrt(0.01 / dm);
  double h = fmin(fmin(100.0 * h0, h1), fmin(max_step, span));

  double D[kNumD];
#pragma unroll
  for (int j = 0; j < kNumD; ++j) D[j] = 0.0;
  D[0] = y;
  D[1] = act ? f0 * h : 0.0;
  rec.init(D, h, lane);
  int order = 1, n_equal = 0;
  double t = t0;
  publish_y(N, L, y, lane);
  jac_eval_for<Lin, G>(N, p, L, rc, lane);
  lin.load_jac(N, L, lane);
  st.jac++;
  bool jac_current = true, lu_valid = false;
  double c_lu = 0.0;
  const double newton_tol = fmax(10.0 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));
  int steps = 0;
  double err = 0.0, d = 0.0;

  while (t < t1) {
    if (steps++ > max_steps) return -2;
    const double hmin = 10.0 * fabs(nextafter(t, INFINITY) - t);
    if (h > max_step) { rec.change(max_step / h); change_D(L, lane, D, order, max_step / h); h = max_step; n_equal = 0; lu_valid = false; }
    if (h < hmin) { rec.change(hmin / h); change_D(L, lane, D, order, hmin / h); h = hmin; n_equal = 0; lu_valid = false; }
    if (t + h > t1) { rec.change((t1 - t) / h); change_D(L, lane, D, order, (t1 - t) / h); h = t1 - t; n_equal = 0; lu_valid = false; }

    double ynew = 0.0, c_acc = 0.0;
    for (;;) {  // attempt the step until accepted
      const double tn = fmin(t + h, t1);
      double ypred = 0.0, psi = 0.0;
#pragma unroll
      for (int j = 0; j < kNumD; ++j)
        if (j <= order) ypred += D[j];
#pragma unroll
      for (int j = 1; j < kNumD; ++j)
        if (j <= order) psi += gam(j) * D[j];
      psi /= alp(order);
      scale = atol + rtol * fabs(ypred);
      const double c = h / alp(order);
      c_acc = c;
      bool converged = false;
      for (;;) {
        if (!lu_valid || c != c_lu) {
          lin.factor(N, L, c, lane);
          st.lu++;
          lu_valid = true;
          c_lu = c;
        }
        d = 0.0;
        ynew = ypred;
        double dy_old = -1.0, rate = -1.0;
        converged = false;
        for (int it = 0; it < 4; ++it) {
          publish_y(N, L, ynew, lane);
          double f = rhs();
          st.rhs++;
          if (!g.all_finite(act ? f : 0.0)) break;
          double dy = act ? c * f - psi - d : 0.0;
          dy = lin.solve(N, L, dy, lane);
          if (!act) dy = 0.0;
          const double dyn = rms_g(g, act ? dy / scale : 0.0, ns);
          if (dy_old >= 0.0) rate = dyn / dy_old;
          // rate^(4 - it) for it = 1..3 by products: pow would cost ~100
          // instructions a Newton iteration for an integer power
          const double rp = it == 1 ? rate * rate * rate : (it == 2 ? rate * rate : rate);
          if (rate >= 0.0 && (rate >= 1.0 || rp / (1.0 - rate) * dyn > newton_tol))
            break;
          ynew += dy;
          d += dy;
          if (dyn == 0.0 || (rate >= 0.0 && rate / (1.0 - rate) * dyn < newton_tol)) {
            converged = true;
            break;
          }
          dy_old = dyn;
        }
        if (converged) break;
        if (!jac_current) {
          publish_y(N, L, ypred, lane);
          jac_eval_for<Lin, G>(N, p, L, rc, lane);
          lin.load_jac(N, L, lane);
          st.jac++;
          jac_current = true;
          lu_valid = false;
          continue;
        }
        break;
      }
      if (!converged) {
        h *= 0.5;
        if (!(h >= hmin)) return -1;  // also ends the span on a NaN step size
        rec.change(0.5);
        change_D(L, lane, D, order, 0.5);
        n_equal = 0;
        lu_valid = false;
        st.rejects++;
        continue;
      }
      scale = atol + rtol * fabs(ynew);
      err = act ? ec(order) * d : 0.0;
      const double en = rms_g(g, act ? err / scale : 0.0, ns);
      if (en > 1.0) {
        double factor = fmax(0.2, 0.9 * pow(en, -1.0 / (order + 1)));
        h *= factor;
        if (!(h > 0.0)) return -1;  // step size underflow or NaN
        rec.change(factor);
        change_D(L, lane, D, order, factor);
        n_equal = 0;
        lu_valid = false;
        st.rejects++;
        continue;
      }
      t = tn;
      break;
    }
    st.steps++;
    jac_current = false;
    // update the differences (D[order+2] = d - D[order+1]; D[order+1] = d)
#pragma unroll
    for (int j = kNumD - 1; j >= 2; --j)
      if (j == order + 2) D[j] = d - D[j - 1];
#pragma unroll
    for (int j = kNumD - 1; j >= 1; --j)
      if (j == order + 1) D[j] = d;
#pragma unroll
    for (int j = kNumD - 2; j >= 0; --j)
      if (j <= order) D[j] += D[j + 1];
    rec.accept(order, c_acc, D, lane);
    n_equal++;
    if (t >= t1) break;
    if (n_equal < order + 1) continue;
    // order selection
    scale = atol + rtol * fabs(D[0]);
    const double e0 = rms_g(g, act ? err / scale : 0.0, ns);
    double em = INFINITY, ep = INFINITY;
    if (order > 1) {
      double dm1 = 0.0;
#pragma unroll
      for (int j = 0; j < kNumD; ++j)
        if (j == order) dm1 = D[j];
      em = rms_g(g, act ? ec(order - 1) * dm1 / scale : 0.0, ns);
    }
    if (order < kMaxOrder) {
      double dp2 = 0.0;
#pragma unroll
      for (int j = 0; j < kNumD; ++j)
        if (j == order + 2) dp2 = D[j];
      ep = rms_g(g, act ? ec(order + 1) * dp2 / scale : 0.0, ns);
    }
    // the three candidate factors e^(-1/(q+1)) in one pow: lanes 0, 1, 2
    const double ev = lane == 0 ? em : (lane == 1 ? e0 : ep);
    const double ex = -1.0 / (double)(order + (lane == 0 ? 0 : (lane == 1 ? 1 : 2)));
    const double fv = ev > 0.0 ? pow(ev, ex) : INFINITY;
    double fm = order > 1 ? g.bc(fv, 0) : 0.0;
    double fz = g.bc(fv, 1);
    double fp = order < kMaxOrder ? g.bc(fv, 2) : 0.0;
    int dord = 0;
    double best = fz;
    if (fm > best) { best = fm; dord = -1; }
    if (fp > best) { best = fp; dord = 1; }
    order += dord;
    const double factor = fmin(10.0, 0.9 * best);
    h *= factor;
    rec.change(factor);
    change_D(L, lane, D, order, factor);
    n_equal = 0;
    lu_valid = false;
  }
  y = D[0];
  return 0;
}

template <class Lin, class G = Wave1>
__device__ __forceinline__ int bdf_integrate(const NetDev &N, const CellParams &p, const Lds &L,
                                             Lin &lin, const RxCache &rc, int lane, double &y,
                                             double t0, double t1, double rtol, double atol,
                                             double max_step, int max_steps, BdfStats &st) {
  NoRec rec;
  return bdf_integrate_rec<Lin, G, NoRec>(N, p, L, lin, rc, lane, y, t0, t1, rtol, atol,
                                          max_step, max_steps, st, rec);
}

// H+/e- steady-state correction (solver.py:227-281) and conservation
// (solver.py:345-388); mode 0 = upstream clipped scales, 1 = exact projection.
template <class G = Wave1>
__device__ void post_solve(const NetDev &N, const CellParams &p, int lane, double &y, int mode,
                           double x_c_total, double x_o_total, double *red = nullptr) {
  const G g{red};
  const int ns = N.ns;
  const bool act = lane < ns;
  const int iH = N.roles[0], iH2 = N.roles[1], iHp = N.roles[2], iCp = N.roles[3],
            iC = N.roles[4], iCO = N.roles[5], iO = N.roles[6], ie = N.e_idx;
  const double *wH = N.post_w, *wC = N.post_w + ns, *wO = N.post_w + 2 * ns,
               *wI = N.post_w + 3 * ns;
  {
    const double xH = g.bc(y, iH), xCp = g.bc(y, iCp);
    const double R = p.zeta * 0.46 * xH;
    const double Tc = fmax(p.T, 10.0);
    const double ae = 2.753e-14 * pow(Tc / 300.0, -0.745) + 1.0e-14 * sqrt(Tc / 100.0);
    const double a = ae * p.nH, b = ae * p.nH * xCp, c = -R;
    if (a > 0.0) {
      const double disc = b * b - 4.0 * a * c;
      if (disc >= 0.0 && xCp < 1e-6) {
        const double xhp = fmax((-b + sqrt(disc)) / (2.0 * a), 1e-30);
        if (lane == iHp) y = xhp;
        const double npos = g.sum(act ? wI[lane] * y : 0.0);
        if (lane == ie) y = fmax(npos, 1e-30);
      }
    }
  }
  {
    const double Hmaj = g.bc(y, iH) + 2.0 * g.bc(y, iH2);
    const double Hmin = g.sum(act ? wH[lane] * y : 0.0);
    if (Hmaj > 0.0) {
      double s = mode == 0 ? fmin(fmax(fmax(1.0 - Hmin, 0.9) / Hmaj, 0.8), 1.2) : (1.0 - Hmin) / Hmaj;
      if (lane == iH || lane == iH2) y *= s;
    }
    const double Cmaj = g.bc(y, iCp) + g.bc(y, iC) + g.bc(y, iCO);
    const double Cmin = g.sum(act ? wC[lane] * y : 0.0);
    if (Cmaj > 0.0) {
      double s = mode == 0
                     ? fmin(fmax(fmax(x_c_total - Cmin, x_c_total * 0.9) / Cmaj, 0.8), 1.2)
                     : (x_c_total - Cmin) / Cmaj;
      if (lane == iCp || lane == iC || lane == iCO) y *= s;
    }
    const double Omol = g.sum(act ? wO[lane] * y : 0.0);
    if (lane == iO) y = fmax(x_o_total - Omol, 1e-30);
    const double npos = g.sum(act ? wI[lane] * y : 0.0);
    if (lane == ie) y = fmax(npos, 1e-30);
    y = fmax(y, 1e-30);
  }
}

__device__ __forceinline__ CellParams load_params(int64_t c, const double *nH, const double *T,
                                                  const double *G0, const double *AV,
                                                  const double *zeta, const double *fH2,
                                                  const double *fCO) {
  CellParams p;
  p.nH = nH[c]; p.T = T[c]; p.G0 = G0[c]; p.AV = AV[c];
  p.zeta = zeta[c]; p.fH2 = fH2[c]; p.fCO = fCO[c];
  p.g0_sqrt_t = p.G0 * sqrt(p.T);
  return p;
}

// The steady-state solve of one cell by one workgroup (solver.py:145-225):
// log-spaced checkpoint spans of the BDF, the early-stop test, the post-solve
// corrections, then the outputs.  The RHS stoichiometry lists must be staged
// in LDS (one-wavefront forms).
template <class Lin, class G>
__device__ __forceinline__ void solve_cell(const NetDev &N, const Lds &L, int64_t cell,
                                           int64_t n, double *__restrict__ y,
                                           const double *__restrict__ nH,
                                           const double *__restrict__ T,
                                           const double *__restrict__ G0,
                                           const double *__restrict__ AV,
                                           const double *__restrict__ zeta,
                                           const double *__restrict__ fH2,
                                           const double *__restrict__ fCO,
                                           const p3d_solve_opts &o, int32_t *__restrict__ conv_out,
                                           int32_t *__restrict__ stats_out) {
  const int lane = threadIdx.x;
  const G g{L.red};
  const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
  const bool act = lane < N.ns;
  double yc = act ? y[(int64_t)lane * n + cell] : 0.0;
  rate_coeffs<G::W>(N, p, L, lane, o.rate_fp32 != 0);
  __syncthreads();
  RxCache rc;
  if constexpr (G::W == 64) rc.load(N, L, lane);

  double tmax = o.t_max;
  if (tmax <= 0.0) tmax = fmin(3e17, fmax(1e16, 3e17 / fmax(p.nH, 1.0)));
  const double l0 = log10(o.t_start), l1 = log10(tmax);
  BdfStats st = {0, 0, 0, 0, 0};
  Lin lin;
  int conv = 0;
  const int nk = o.n_checkpoints;
  for (int i = 0; i < nk - 1; ++i) {
    double ta = (i == 0) ? o.t_start : pow(10.0, l0 + (l1 - l0) * i / (nk - 1));
    double tb = (i + 1 == nk - 1) ? tmax : pow(10.0, l0 + (l1 - l0) * (i + 1) / (nk - 1));
    double yn = yc;
    int status = bdf_integrate<Lin, G>(N, p, L, lin, rc, lane, yn, ta, tb, o.rtol, o.atol,
                                       tb - ta, o.max_steps, st);
    if (status != 0) continue;  // a failed span leaves the state unchanged (solver.py:216-217)
    yn = fmax(yn, 1e-30);
    if (o.early_stop) {
      double rel = -1.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int k = N.roles[q == 0 ? 0 : q == 1 ? 1 : q == 2 ? 3 : q == 3 ? 4 : q == 4 ? 5 : 6];
        const double old = g.bc(yc, k), nw = g.bc(yn, k);
        if (old > 1e-12) rel = fmax(rel, fabs(nw - old) / fmax(old, 1e-20));
      }
      if (rel >= 0.0 && rel < o.stop_rtol && i > 3) {
        conv = 1;
        yc = yn;
        break;
      }
    }
    yc = yn;
  }
  if (o.post_mode >= 0)
    post_solve<G>(N, p, lane, yc, o.post_mode, o.x_c_total, o.x_o_total, L.red);
  if (act) y[(int64_t)lane * n + cell] = yc;
  if (lane == 0) {
    if (conv_out) conv_out[cell] = conv;
    if (stats_out) {
      stats_out[4 * cell + 0] = st.steps;
      stats_out[4 * cell + 1] = st.lu;
      stats_out[4 * cell + 2] = st.rhs;
      stats_out[4 * cell + 3] = st.jac;
    }
  }
}

template <class Lin, class G = Wave1>
__global__ void __launch_bounds__(G::W)
chem_solve_kernel(NetDev N, int64_t n, double *__restrict__ y, const double *__restrict__ nH,
                  const double *__restrict__ T, const double *__restrict__ G0,
                  const double *__restrict__ AV, const double *__restrict__ zeta,
                  const double *__restrict__ fH2, const double *__restrict__ fCO,
                  p3d_solve_opts o, int32_t *__restrict__ conv_out,
                  int32_t *__restrict__ stats_out, const int64_t *__restrict__ cell_list,
                  const unsigned *__restrict__ list_len) {
  extern __shared__ double smem[];
  // with a cell list (cells handed over by the lane kernel), workgroup b
  // solves cell_list[b] if b < *list_len
  if (cell_list && blockIdx.x >= *list_len) return;
  const int64_t cell = cell_list ? cell_list[blockIdx.x] : (int64_t)blockIdx.x;
  Lds L;
  if constexpr (G::W == 64) L = carve(smem, N);
  else L = carve_wide(smem, N);
  if constexpr (G::W == 64) stage_tables(N, L, threadIdx.x);
  if constexpr (std::is_same<Lin, LinSparseG<G>>::value) Lin::stage(N, L, threadIdx.x);
  solve_cell<Lin, G>(N, L, cell, n, y, nH, T, G0, AV, zeta, fH2, fCO, o, conv_out, stats_out);
}

// Persistent form (variant 6): each resident wavefront takes cells from the
// device work counter one at a time until the batch is exhausted, so cells
// whose step counts differ by orders of magnitude share the chip without
// idle workgroups.  One cell per wavefront, lane = species, M in registers.
__global__ void __launch_bounds__(64) queue_reset(unsigned long long *ctr) {
  if (threadIdx.x == 0) *ctr = 0ull;
}

//
// Two wavefronts per SIMD (8 per CU, also the LDS limit at ~20 KB per cell):
// the compiler then fits the kernel in 256 registers and spills the rest to
// scratch.  A lone wavefront leaves the SIMD idle on every dependent FP64
// and readlane latency; the second one fills them.  On the 64^3 workload
// this solves the same cells bit for bit in 2.94 s against 4.12 s with one
// wavefront per SIMD (410 registers, no scratch).
#ifndef P3D_QUEUE_WPE
#define P3D_QUEUE_WPE 2
#endif
template <class Lin>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(P3D_QUEUE_WPE)))
chem_queue_kernel(NetDev N, int64_t n, double *__restrict__ y, const double *__restrict__ nH,
                  const double *__restrict__ T, const double *__restrict__ G0,
                  const double *__restrict__ AV, const double *__restrict__ zeta,
                  const double *__restrict__ fH2, const double *__restrict__ fCO,
                  p3d_solve_opts o, int32_t *__restrict__ conv_out,
                  int32_t *__restrict__ stats_out, unsigned long long *__restrict__ ctr) {
  extern __shared__ double smem[];
  const Lds L = carve_wave(smem, N);
  for (;;) {
    unsigned long long c = 0;
    if (threadIdx.x == 0) c = atomicAdd(ctr, 1ull);
    c = __shfl(c, 0, 64);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)c);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(c >> 32));
    const int64_t cell = (int64_t)(((unsigned long long)hi << 32) | lo);
    if (cell >= n) break;  // scalar branch: every lane leaves together
    solve_cell<Lin, Wave1>(N, L, cell, n, y, nH, T, G0, AV, zeta, fH2, fCO, o, conv_out,
                           stats_out);
    __syncthreads();  // LDS of this cell is rewritten by the next
  }
}

// The upstream fast mode (chemistry/solver.py:164-183), one cell per
// workgroup, lane = species: ten explicit steps y <- max(y + dt f(y), 1e-30)
// with dt limited so that the largest relative change |f| dt / max(|y|,
// 1e-20) is 0.3, then doubled (at most 1e15 s), then the post-solve
// corrections.  Uncontracted, so every step rounds as the numpy expression.
__global__ void __launch_bounds__(64)
chem_fast_kernel(NetDev N, int64_t n, double *__restrict__ y, const double *__restrict__ nH,
                 const double *__restrict__ T, const double *__restrict__ G0,
                 const double *__restrict__ AV, const double *__restrict__ zeta,
                 const double *__restrict__ fH2, const double *__restrict__ fCO,
                 p3d_solve_opts o, int32_t *__restrict__ conv_out,
                 int32_t *__restrict__ stats_out) {
#pragma clang fp contract(off)
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  const Lds L = carve(smem, N);
  stage_tables(N, L, lane);
  const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
  const bool act = lane < N.ns;
  double yc = act ? y[(int64_t)lane * n + cell] : 0.0;
  rate_coeffs<64>(N, p, L, lane, o.rate_fp32 != 0);
  __syncthreads();
  RxCache rc;
  rc.load(N, L, lane);
  const bool cached = N.nr <= 64 * kRxPerLane;
  double dt = 1e12;
  for (int step = 0; step < 10; ++step) {
    publish_y(N, L, yc, lane);
    const double f = cached ? rhs_cached(N, p, L, rc, lane) : rhs_balanced<64>(N, p, L, lane);
    const double mc = act ? fabs(f) * dt / fmax(fabs(yc), 1e-20) : 0.0;
    const double max_change = wave_max(mc);
    if (max_change > 0.3) dt *= 0.3 / max_change;
    if (act) yc = fmax(yc + f * dt, 1e-30);
    dt *= 2.0;
    dt = fmin(dt, 1e15);
  }
  if (o.post_mode >= 0)
    post_solve<Wave1>(N, p, lane, yc, o.post_mode, o.x_c_total, o.x_o_total, L.red);
  if (act) y[(int64_t)lane * n + cell] = yc;
  if (lane == 0) {
    if (conv_out) conv_out[cell] = 0;
    if (stats_out) {
      stats_out[4 * cell + 0] = 10;
      stats_out[4 * cell + 1] = 0;
      stats_out[4 * cell + 2] = 10;
      stats_out[4 * cell + 3] = 0;
    }
  }
}

__global__ void __launch_bounds__(64)
chem_post_kernel(NetDev N, int64_t n, double *__restrict__ y, const double *__restrict__ nH,
                 const double *__restrict__ T, const double *__restrict__ zeta, int mode,
                 double x_c_total, double x_o_total) {
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  CellParams p;
  p.nH = nH[cell]; p.T = T[cell]; p.zeta = zeta[cell];
  p.G0 = 0.0; p.AV = 0.0; p.fH2 = 1.0; p.fCO = 1.0; p.g0_sqrt_t = 0.0;
  double yc = lane < N.ns ? y[(int64_t)lane * n + cell] : 0.0;
  post_solve(N, p, lane, yc, mode, x_c_total, x_o_total);
  if (lane < N.ns) y[(int64_t)lane * n + cell] = yc;
}

__global__ void __launch_bounds__(64)
chem_rhs_kernel(NetDev N, int64_t n, const double *__restrict__ y, const double *__restrict__ nH,
                const double *__restrict__ T, const double *__restrict__ G0,
                const double *__restrict__ AV, const double *__restrict__ zeta,
                const double *__restrict__ fH2, const double *__restrict__ fCO,
                double *__restrict__ dydt) {
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  const Lds L = carve(smem, N);
  const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
  rate_coeffs(N, p, L, lane);
  __syncthreads();
  double yv = lane < N.ns ? y[(int64_t)lane * n + cell] : 0.0;
  publish_y(N, L, yv, lane);
  double f = rhs_lane(N, p, L, lane);
  if (lane < N.ns) dydt[(int64_t)lane * n + cell] = f;
}

__global__ void __launch_bounds__(64)
chem_jac_kernel(NetDev N, int64_t n, const double *__restrict__ y, const double *__restrict__ nH,
                const double *__restrict__ T, const double *__restrict__ G0,
                const double *__restrict__ AV, const double *__restrict__ zeta,
                const double *__restrict__ fH2, const double *__restrict__ fCO,
                double *__restrict__ jac) {
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  const Lds L = carve(smem, N);
  const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
  rate_coeffs(N, p, L, lane);
  __syncthreads();
  double yv = lane < N.ns ? y[(int64_t)lane * n + cell] : 0.0;
  publish_y(N, L, yv, lane);
  jacobian(N, p, L, lane);
  const int ns = N.ns;
  double *out = jac + cell * (int64_t)ns * ns;
  for (int idx = lane; idx < ns * ns; idx += 64) {
    int i = idx / ns, j = idx - i * ns;
    out[idx] = L.J[i * L.ld + j];
  }
}

// ===========================================================================
// Charge-neutral steady state and its adjoint.
//
// The network conserves the nuclei of every element but not charge (PAH
// recombinations move charge into the untracked PAH- pseudo-species), so the
// raw rate equations have no fixed point: free electrons accumulate.  The
// upstream solve closes charge after integrating, with e- = sum of the
// cations (chemistry/solver.py:382-386).  The steady state here is the
// algebraic state that the closure makes exact: every species balance is
// zero except
//   - per element, the row of its pivot species (the most abundant species
//     carrying that element alone), replaced by the element's nuclei total;
//   - the electron row, replaced by e- = sum of the cations;
//   - species at or below 1e-28 (at the 1e-30 floor of the rate equations),
//     held fixed.
// chem_polish_kernel reaches it from an integrated state by pseudo-transient
// continuation (implicit Euler steps of the balances with a growing pseudo
// time step, ending in plain Newton); chem_adjoint_kernel
// differentiates an objective at it with respect to every log10 k_r, ln G0,
// ln zeta and ln n_H by one transposed solve (the implicit function theorem).
// ===========================================================================
constexpr int kMaxElem = 8;
// Species at or below this are held: they sit at the rate equations' 1e-30
// floor (their own balance has no root there); the margin keeps the choice
// the same for integrations that differ by rounding.
constexpr double kFrozenBelow = 1e-28;

struct SteadyCons {  // kernel argument
  int n_elem;
  signed char nuclei[kMaxElem][kMaxSpecies];
  signed char cation[kMaxSpecies];
};

struct SteadyRows {
  int pivot_of;  // element whose constraint replaces this lane's row, or -1
  bool frozen;   // species at the floor: row held fixed
};

// Row roles of lane `lane`'s species at abundance x (all lanes participate).
__device__ SteadyRows steady_rows(const NetDev &N, const SteadyCons &K, double x, int lane) {
  SteadyRows r;
  r.frozen = lane < N.ns && !(x > kFrozenBelow);
  r.pivot_of = -1;
  int carried = 0;
  for (int e = 0; e < K.n_elem; ++e) carried += (lane < N.ns && K.nuclei[e][lane] > 0) ? 1 : 0;
  for (int e = 0; e < K.n_elem; ++e) {
    const bool cand = lane < N.ns && carried == 1 && K.nuclei[e][lane] > 0 && !r.frozen &&
                      lane != N.e_idx;
    const double v = cand ? (double)K.nuclei[e][lane] * x : -1.0;
    const int best = wave_argmax(v, lane);
    const double bv = bcast(v, best);
    if (bv > 0.0 && lane == best) r.pivot_of = e;
  }
  return r;
}

// Writes row `lane` of the constrained Jacobian into L.M, transposed when
// `transpose` (L.J holds the network Jacobian); the balance rows carry the
// pseudo-time term -1/dt on the diagonal (0 for the plain Jacobian).
__device__ void steady_matrix(const NetDev &N, const SteadyCons &K, const Lds &L,
                              const SteadyRows &r, bool transpose, int lane,
                              double inv_dt = 0.0) {
  const int n = N.ns, ld = L.ld;
  if (lane < n) {
    for (int j = 0; j < n; ++j) {
      double v;
      if (r.frozen) v = j == lane ? 1.0 : 0.0;
      else if (lane == N.e_idx) v = (j == lane ? 1.0 : 0.0) - (double)K.cation[j];
      else if (r.pivot_of >= 0) v = (double)K.nuclei[r.pivot_of][j];
      else v = L.J[lane * ld + j] - (j == lane ? inv_dt : 0.0);
      if (transpose) L.M[j * ld + lane] = v;
      else L.M[lane * ld + j] = v;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ double lane_sum(const NetDev &N, double v, int lane) {
  return wave_sum(lane < N.ns ? v : 0.0);
}

__global__ void __launch_bounds__(64)
chem_polish_kernel(NetDev N, SteadyCons K, int64_t n, double *__restrict__ y,
                   const double *__restrict__ nH, const double *__restrict__ T,
                   const double *__restrict__ G0, const double *__restrict__ AV,
                   const double *__restrict__ zeta, const double *__restrict__ fH2,
                   const double *__restrict__ fCO, const double *__restrict__ y_tot,
                   int max_iter, int32_t *__restrict__ iters) {
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  const Lds L = carve(smem, N);
  const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
  rate_coeffs(N, p, L, lane);
  __syncthreads();
  double x = lane < N.ns ? y[(int64_t)lane * n + cell] : 0.0;
  // element totals (of the reference state when given), and the electron closure
  const double xt = (y_tot && lane < N.ns) ? y_tot[(int64_t)lane * n + cell] : x;
  double tot[kMaxElem];
  for (int e = 0; e < K.n_elem; ++e)
    tot[e] = lane_sum(N, lane < N.ns ? (double)K.nuclei[e][lane] * xt : 0.0, lane);
  const double cat0 = lane_sum(N, lane < N.ns ? (double)K.cation[lane] * x : 0.0, lane);
  if (lane == N.e_idx) x = cat0;
  const SteadyRows rows = steady_rows(N, K, x, lane);
  // pseudo-transient continuation: implicit Euler steps of the balances with
  // a pseudo time step from 1e10 s, scaled after each step so the largest
  // relative change aims at 0.5, or 0.1 after 100 iterations (x0.1 ... x10;
  // x0.1 when positivity damps the step), up to 1e30 s where the iteration
  // is plain Newton
  double dt = 1e10;
  int it = 0, done = 0;
  for (; it < max_iter && !done; ++it) {
    publish_y(N, L, x, lane);
    const double f = rhs_lane(N, p, L, lane);
    jacobian(N, p, L, lane);
    // wave reductions run on every lane, outside the per-lane selection
    const double cat = lane_sum(N, lane < N.ns ? (double)K.cation[lane] * x : 0.0, lane);
    double F;
    if (rows.frozen) F = 0.0;
    else if (lane == N.e_idx) F = x - cat;
    else F = f;
    // element rows (every lane takes part in the sums)
    for (int e = 0; e < K.n_elem; ++e) {
      const double c = lane_sum(N, lane < N.ns ? (double)K.nuclei[e][lane] * x : 0.0, lane) - tot[e];
      if (rows.pivot_of == e) F = c;
    }
    steady_matrix(N, K, L, rows, false, lane, 1.0 / dt);
    lu_factor_inplace(N, L, lane);
    const double d = lu_solve(N, L, lane < N.ns ? -F : 0.0, lane);
    const double lim = (lane < N.ns && d < 0.0) ? -0.9 * x / d : INFINITY;
    const double a = fmin(1.0, -wave_max(-lim));
    if (lane < N.ns) x += a * d;
    const double rel = wave_max(lane < N.ns ? fabs(d) / fmax(fabs(x), 1e-30) : 0.0);
    done = (a == 1.0 && dt >= 1e30 && rel < 1e-10) ? 1 : 0;
    const double target = it < 100 ? 0.5 : 0.1;  // gentler for cold dense cells
    const double grow = a < 1.0 ? 0.1 : fmin(fmax(target / fmax(rel, 1e-300), 0.1), 10.0);
    dt = fmin(fmax(dt * grow, 1.0), 1e30);
  }
  if (lane < N.ns) y[(int64_t)lane * n + cell] = x;
  if (lane == 0 && iters) iters[cell] = done ? it : -1;
}

struct CiiLine {
  double freq, A, Tul, gu, gl;
};

// The differentiation objective of one cell at abundances x (lane's
// species): the [CII] 158 um cooling (solver_3d.py:956-965) weighted by
// w_cii plus the CO column density n_H x_CO weighted by w_co.  g: its
// derivative with respect to the lane's abundance; explicit_n: its explicit
// derivative with respect to ln n_H (abundances held).
struct CellObjective {
  double g, lam, co, explicit_n;
};
__device__ CellObjective cell_objective(const NetDev &N, const CellParams &p, const CiiLine &cl,
                                        double w_cii, double w_co, double x, int lane) {
#pragma clang fp contract(off)
  const int iH = N.roles[0], iH2 = N.roles[1], iCp = N.roles[3], iCO = N.roles[5];
  const double xcp = bcast(x, iCp), xe = bcast(x, N.e_idx), xh = bcast(x, iH),
               xh2 = bcast(x, iH2), xco = bcast(x, iCO);
  const double Tk = p.T, nh = p.nH;
  const double ge = 8.63e-6 / (cl.gl * sqrt(Tk)) * 2.15 * exp(-cl.Tul / Tk);
  const double gh = 8.0e-10 * pow(Tk / 100.0, 0.07);
  const double gh2 = 4.9e-10 * pow(Tk / 100.0, 0.12);
  const double q = ge * (xe * nh) + gh * (xh * nh) + gh2 * (xh2 * nh);
  const double ar = (cl.gu / cl.gl) * exp(-cl.Tul / Tk);
  const double frac = ar * q / (ar * q + cl.A + q);
  const double eph = cl.A * 6.62607015e-27 * cl.freq;
  const double lam = xcp * nh * frac * eph;
  const double dfq = ar * cl.A / (((1.0 + ar) * q + cl.A) * ((1.0 + ar) * q + cl.A));
  const double base = xcp * nh * eph * dfq;
  double g = 0.0;
  if (lane == iCp) g += w_cii * (nh * frac * eph);
  if (lane == N.e_idx) g += w_cii * (base * ge * nh);
  if (lane == iH) g += w_cii * (base * gh * nh);
  if (lane == iH2) g += w_cii * (base * gh2 * nh);
  if (lane == iCO) g += w_co * nh;
  CellObjective o;
  o.g = g;
  o.lam = lam;
  o.co = nh * xco;
  o.explicit_n = w_cii * (lam + base * q) + w_co * (nh * xco);
  return o;
}

__global__ void __launch_bounds__(64)
chem_adjoint_kernel(NetDev N, SteadyCons K, int64_t n, const double *__restrict__ y,
                    const double *__restrict__ nH, const double *__restrict__ T,
                    const double *__restrict__ G0, const double *__restrict__ AV,
                    const double *__restrict__ zeta, const double *__restrict__ fH2,
                    const double *__restrict__ fCO, CiiLine cl, double w_cii, double w_co,
                    double *__restrict__ obj, double *__restrict__ grad) {
#pragma clang fp contract(off)
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  const Lds L = carve(smem, N);
  const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
  rate_coeffs(N, p, L, lane);
  __syncthreads();
  const double x = lane < N.ns ? y[(int64_t)lane * n + cell] : 0.0;
  publish_y(N, L, x, lane);
  jacobian(N, p, L, lane);
  const SteadyRows rows = steady_rows(N, K, x, lane);
  steady_matrix(N, K, L, rows, true, lane);
  const CellObjective ob = cell_objective(N, p, cl, w_cii, w_co, x, lane);
  lu_factor_inplace(N, L, lane);
  const double lmb = lu_solve(N, L, lane < N.ns ? ob.g : 0.0, lane);
  // multipliers of the rows that move with the parameters (the species
  // balances); the constraint and frozen rows do not
  const bool dyn = lane < N.ns && !rows.frozen && rows.pivot_of < 0 && lane != N.e_idx;
  __syncthreads();  // the factors are no longer read
  if (lane < N.ns) L.M[lane] = dyn ? lmb : 0.0;
  eval_reactions(N, p, L, lane, false);  // rates into part[2r] (syncs)
  // PAH charge parameter of the closed form (solver.py:74-87)
  const double yef = fmax(L.ys[N.e_idx], 1e-10);
  const double ne = yef * p.nH;
  const double psi = (p.G0 > 0.0 && ne > 0.0) ? p.G0 * sqrt(p.T) / ne : -1.0;
  const double fp = psi > 0.0 ? psi / (psi + 500.0) : 0.0;
  double sG = 0.0, sZ = 0.0, sN = 0.0;
  double *gout = grad + cell * (int64_t)(N.nr + 3);
  for (int r = lane; r < N.nr; r += 64) {
    double w = 0.0;  // sum over species of multiplier x net stoichiometry
    for (int s2 = 0; s2 < N.ns; ++s2) {
      const double l2 = L.M[s2];
      if (l2 == 0.0) continue;
      for (int e = N.rhs_off[s2]; e < N.rhs_off[s2 + 1]; ++e)
        if (N.rhs_rxn[e] == r) w += N.rhs_sgn[e] * l2;
    }
    const double rate = L.part[2 * r];
    const int t = N.rtype[r];
    const bool pah = N.react[2 * r + 1] == P3D_SP_PAH_MINUS;
    gout[r] = -(2.302585092994046 * rate) * w;
    const double wr = w * rate;
    const bool photo = t == P3D_RX_PHOTO || t == P3D_RX_PHOTO_FH2 || t == P3D_RX_PHOTO_FCO;
    sG += wr * ((photo ? 1.0 : 0.0) - (pah && t == P3D_RX_TWO_BODY ? fp : 0.0));
    sZ += wr * (t == P3D_RX_CR ? 1.0 : 0.0);
    sN += wr * (t == P3D_RX_TWO_BODY ? 1.0 + (pah ? fp : 0.0) : (t == P3D_RX_GRAIN ? 1.0 : 0.0));
  }
  sG = wave_sum(sG);
  sZ = wave_sum(sZ);
  sN = wave_sum(sN);
  if (lane == 0) {
    gout[N.nr] = -sG;
    gout[N.nr + 1] = -sZ;
    gout[N.nr + 2] = ob.explicit_n - sN;
    obj[2 * cell] = ob.lam;
    obj[2 * cell + 1] = ob.co;
  }
}

// ===========================================================================
// Discrete adjoint of the implicit steps (bundled network, one cell per
// wavefront).  The forward map is the BDF integration of a cell from its
// initial abundances over one span (chemistry/solver.py:185-217) with the
// step sequence held fixed: the step sizes, orders and step-size changes the
// integration chose are data, the implicit steps are differentiated exactly.
//
// The taped forward (chem_tape_kernel) records per accepted step its
// c = h / alpha_k, its order k and the product F of the step-size changes
// applied to the differences before it (changes compose: rescaling by f1
// then f2 is rescaling by f1 f2), plus the differences D every 16 steps.
// The adjoint (chem_tape_adjoint_kernel) walks the 16-step segments from the
// last: it recomputes the segment from its checkpoint (Newton on each
// implicit step with the recorded c, k, F), stores per step the factors of
// I - c J(y_n+1) in the column layout, then runs the segment backwards:
//   d' = adjoint of the difference update, lambda = (I - c J)^-T d',
//   differences' += lambda - d' (predictor) and -gamma_j / alpha_k lambda
//   (psi), parameters' += c (df/dtheta)^T lambda, then the adjoint of the
//   step-size change (D' <- RU D' on rows 0..k);
// and ends with D[1] = h0 f(y0): parameters' += h0 (df/dtheta(y0))^T D'_1.
// ===========================================================================
constexpr int kTapeSeg = 16;     // steps per checkpoint segment
constexpr int kTapeCol = 33;     // doubles per lane per stored step: 31 factors, 1/u, y
constexpr double kLn10 = 2.302585092994046;

__constant__ unsigned c_wave_tum[64] = P3D_WAVE_TUMASK;
__constant__ unsigned c_wave_tlm[64] = P3D_WAVE_TLMASK;
__constant__ int c_wave_tlev[64] = P3D_WAVE_TLEVELS;

struct TapeDev {   // kernel argument: the tape of n cells
  double *rec;     // n x cap x 3: per accepted step {c, F, k}
  double *ck;      // n x (cap / 16 + 1) x kNumD x 32: differences every 16 steps
  double *hinit;   // n: step size of D[1] = h f(y0)
  int32_t *nsteps; // n: accepted steps; -1 failed integration, -2 tape too short
  int cap;         // steps per cell (a multiple of 16)
};

struct TapeRec {
  double *rec, *ck;
  int cap, n;
  double F, h0;
  bool over;
  __device__ __forceinline__ void ckpt(const double (&D)[kNumD], int lane) {
    double *c = ck + (size_t)(n / kTapeSeg) * kNumD * 32;
    if (lane < 32) {
#pragma unroll
      for (int j = 0; j < kNumD; ++j) c[j * 32 + lane] = D[j];
    }
  }
  __device__ __forceinline__ void init(const double (&D)[kNumD], double h, int lane) {
    F = 1.0;
    n = 0;
    over = false;
    h0 = h;
    ckpt(D, lane);
  }
  __device__ __forceinline__ void change(double f) { F *= f; }
  __device__ __forceinline__ void accept(int order, double c, const double (&D)[kNumD],
                                         int lane) {
    if (n < cap) {
      const double v = lane == 0 ? c : (lane == 1 ? F : (double)order);
      if (lane < 3) rec[3 * n + lane] = v;
    } else {
      over = true;
    }
    F = 1.0;
    ++n;
    if (n % kTapeSeg == 0 && n < cap) ckpt(D, lane);
  }
};

__device__ __forceinline__ int64_t next_cell(unsigned long long *ctr) {
  unsigned long long c = 0;
  if (threadIdx.x == 0) c = atomicAdd(ctr, 1ull);
  c = __shfl(c, 0, 64);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)c);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(c >> 32));
  return (int64_t)(((unsigned long long)hi << 32) | lo);
}

// Taped forward: the one-span integration of chem_queue_kernel (same
// solve, same results) with the step sequence recorded.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
chem_tape_kernel(NetDev N, int64_t n, double *__restrict__ y, const double *__restrict__ nH,
                 const double *__restrict__ T, const double *__restrict__ G0,
                 const double *__restrict__ AV, const double *__restrict__ zeta,
                 const double *__restrict__ fH2, const double *__restrict__ fCO,
                 p3d_solve_opts o, TapeDev TD, int32_t *__restrict__ stats_out,
                 unsigned long long *__restrict__ ctr) {
  extern __shared__ double smem[];
  const Lds L = carve_wave(smem, N);
  const int lane = threadIdx.x;
  for (;;) {
    const int64_t cell = next_cell(ctr);
    if (cell >= n) break;  // scalar branch: every lane leaves together
    const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
    const bool act = lane < N.ns;
    double yc = act ? y[(int64_t)lane * n + cell] : 0.0;
    rate_coeffs<64>(N, p, L, lane, o.rate_fp32 != 0);
    __syncthreads();
    RxCache rc;
    rc.load(N, L, lane);
    LinStatic lin;
    double tmax = o.t_max;
    if (tmax <= 0.0) tmax = fmin(3e17, fmax(1e16, 3e17 / fmax(p.nH, 1.0)));
    TapeRec rec;
    rec.rec = TD.rec + (size_t)cell * TD.cap * 3;
    rec.ck = TD.ck + (size_t)cell * (TD.cap / kTapeSeg + 1) * kNumD * 32;
    rec.cap = TD.cap;
    BdfStats st = {0, 0, 0, 0, 0};
    double yn = yc;
    const int status = bdf_integrate_rec<LinStatic, Wave1, TapeRec>(
        N, p, L, lin, rc, lane, yn, o.t_start, tmax, o.rtol, o.atol, tmax - o.t_start,
        o.max_steps, st, rec);
    if (status == 0) yc = fmax(yn, 1e-30);
    if (act) y[(int64_t)lane * n + cell] = yc;
    if (lane == 0) {
      TD.nsteps[cell] = status != 0 ? -1 : (rec.over ? -2 : rec.n);
      TD.hinit[cell] = rec.h0;
      if (stats_out) {
        stats_out[4 * cell + 0] = st.steps;
        stats_out[4 * cell + 1] = st.lu;
        stats_out[4 * cell + 2] = st.rhs;
        stats_out[4 * cell + 3] = st.jac;
      }
    }
    __syncthreads();  // LDS of this cell is rewritten by the next
  }
}

// Adjoint of change_D: D' <- RU D' on rows 0..order (change_D applies RU^T).
__device__ void change_D_adj(const Lds &L, int lane, double (&Db)[kNumD], int order,
                             double factor) {
  if (lane < 36) {
    const int i = lane / 6, j = lane % 6;
    double acc = 0.0;
    if (i <= order && j <= order) {
      for (int m = 0; m <= order; ++m) {
        double r = 1.0, u = 1.0;
        for (int q = 1; q <= i; ++q) r *= ((double)(q - 1) - factor * (double)m) / (double)q;
        for (int q = 1; q <= m; ++q) u *= ((double)(q - 1) - (double)j) / (double)q;
        acc += r * u;
      }
    }
    L.ru[lane] = acc;
  }
  __syncthreads();
  double Dn[kMaxOrder + 1];
#pragma unroll
  for (int i = 0; i <= kMaxOrder; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j <= kMaxOrder; ++j) acc += L.ru[6 * i + j] * Db[j];
    Dn[i] = acc;
    __asm__ volatile("" ::: "memory");
  }
#pragma unroll
  for (int i = 0; i <= kMaxOrder; ++i)
    if (i <= order) Db[i] = Dn[i];
  __syncthreads();
}

// w (df/dtheta(y))^T lambda accumulated per lane: reaction lane + 64 q into
// gk[q] (d/dlog10 k), and the lane's shares of d/dln G0, d/dln zeta and
// d/dln n_H (the rates' explicit dependence, the PAH charge balance's
// through psi = G0 sqrt(T) / n_e included).  Requires ys published (y) and
// psum[s] = lambda_s (0 past the species).
__device__ void param_adjoint(const NetDev &N, const CellParams &p, const Lds &L, int lane,
                              double w, double (&gk)[kRxPerLane], double &sG, double &sZ,
                              double &sN) {
  eval_reactions<64>(N, p, L, lane, false);  // rates into part[2r] (syncs)
  const double yef = fmax(L.ys[N.e_idx], 1e-10);
  const double ne = yef * p.nH;
  const double psi = (p.G0 > 0.0 && ne > 0.0) ? p.g0_sqrt_t / ne : -1.0;
  const double fp = psi > 0.0 ? psi / (psi + 500.0) : 0.0;
#pragma unroll
  for (int q = 0; q < kRxPerLane; ++q) {
    const int r = lane + 64 * q;
    if (r >= N.nr) continue;
    const int *s5 = N.st5 + 5 * r;
    double wr = 0.0;
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (s5[m] >= 0) wr -= L.psum[s5[m]];
#pragma unroll
    for (int m = 2; m < 5; ++m)
      if (s5[m] >= 0) wr += L.psum[s5[m]];
    const double v = w * (wr * L.part[2 * r]);
    gk[q] += kLn10 * v;
    const int t = N.rtype[r];
    const bool pah = N.react[2 * r + 1] == P3D_SP_PAH_MINUS;
    const bool photo = t == P3D_RX_PHOTO || t == P3D_RX_PHOTO_FH2 || t == P3D_RX_PHOTO_FCO;
    sG += v * ((photo ? 1.0 : 0.0) - (pah && t == P3D_RX_TWO_BODY ? fp : 0.0));
    sZ += v * (t == P3D_RX_CR ? 1.0 : 0.0);
    sN += v * (t == P3D_RX_TWO_BODY ? 1.0 + (pah ? fp : 0.0) : (t == P3D_RX_GRAIN ? 1.0 : 0.0));
  }
  __syncthreads();  // part[] and psum[] are rewritten next
}

// Two wavefronts per SIMD, as the forward solve: 256 VGPRs and 452 B of
// scratch per lane instead of 365 VGPRs at one; the 32^3 cube's adjoint walk
// takes 0.97 s against 1.32 s, the same gradient bit for bit.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
chem_tape_adjoint_kernel(NetDev N, int64_t n, const double *__restrict__ y0,
                         const double *__restrict__ yend, const double *__restrict__ nH,
                         const double *__restrict__ T, const double *__restrict__ G0,
                         const double *__restrict__ AV, const double *__restrict__ zeta,
                         const double *__restrict__ fH2, const double *__restrict__ fCO,
                         p3d_solve_opts o, TapeDev TD, CiiLine cl, double w_cii, double w_co,
                         double *__restrict__ obj, double *__restrict__ grad,
                         int32_t *__restrict__ status_out, double *__restrict__ segbuf,
                         unsigned long long *__restrict__ ctr) {
  extern __shared__ double smem[];
  const Lds L = carve_wave(smem, N);
  const int lane = threadIdx.x;
  const bool act = lane < N.ns;
  const bool row = lane < wave::kNS;
  double *seg = segbuf + (size_t)blockIdx.x * kTapeSeg * kTapeCol * 64;
  const unsigned tlm = row ? c_wave_tlm[lane] : 0u, tum = row ? c_wave_tum[lane] : 0u;
  const int tlev = c_wave_tlev[lane];
  auto gam = [](int k) { return c_gam[k]; };
  auto alp = [](int k) { return (1.0 - c_kappa[k]) * c_gam[k]; };
  const double newton_tol =
      fmax(10.0 * 2.220446049250313e-16 / o.rtol, fmin(0.03, sqrt(o.rtol)));
  for (;;) {
    const int64_t cell = next_cell(ctr);
    if (cell >= n) break;
    const CellParams p = load_params(cell, nH, T, G0, AV, zeta, fH2, fCO);
    rate_coeffs<64>(N, p, L, lane, o.rate_fp32 != 0);
    __syncthreads();
    RxCache rc;
    rc.load(N, L, lane);
    LinStatic lin;
    const double xe = act ? yend[(int64_t)lane * n + cell] : 0.0;
    const CellObjective ob = cell_objective(N, p, cl, w_cii, w_co, xe, lane);
    double Db[kNumD];
#pragma unroll
    for (int j = 0; j < kNumD; ++j) Db[j] = 0.0;
    Db[0] = (act && xe > 1e-30) ? ob.g : 0.0;  // y_end = max(D0, 1e-30)
    double gk[kRxPerLane] = {0.0, 0.0};
    double sG = 0.0, sZ = 0.0, sN = 0.0;
    const int nst = TD.nsteps[cell];
    int failed = nst < 0 ? nst : 0;
    if (nst >= 0) {
      const double *rec = TD.rec + (size_t)cell * TD.cap * 3;
      const double *ck = TD.ck + (size_t)cell * (TD.cap / kTapeSeg + 1) * kNumD * 32;
      const int nseg = (nst + kTapeSeg - 1) / kTapeSeg;
      for (int sg = nseg - 1; sg >= 0; --sg) {
        const int i0 = sg * kTapeSeg, i1 = min(i0 + kTapeSeg, nst);
        double D[kNumD];
#pragma unroll
        for (int j = 0; j < kNumD; ++j) D[j] = act ? ck[(sg * kNumD + j) * 32 + lane] : 0.0;
        // recompute the segment; L.J holds J at the latest state, L.M the
        // factors of I - c_lu J
        double c_lu = -1.0;
        bool jac_ok = false;
        for (int i = i0; i < i1; ++i) {
          const double c = rec[3 * i], F = rec[3 * i + 1];
          const int k = (int)rec[3 * i + 2];
          if (F != 1.0) change_D(L, lane, D, k, F);
          double ypred = 0.0, psi = 0.0;
#pragma unroll
          for (int j = 0; j < kNumD; ++j)
            if (j <= k) ypred += D[j];
#pragma unroll
          for (int j = 1; j < kNumD; ++j)
            if (j <= k) psi += gam(j) * D[j];
          psi /= alp(k);
          const double scale = o.atol + o.rtol * fabs(ypred);
          if (!jac_ok) {
            publish_y(N, L, ypred, lane);
            jacobian_sched<64>(N, p, L, lane);
            jac_ok = true;
            c_lu = -1.0;
          }
          // Newton on the recorded implicit step.  Species at the abundance
          // floor make the equations non-smooth at the atol level, so the
          // iteration may stall there; the recomputed segment is checked
          // against the next checkpoint instead (below).
          double d = 0.0, ynew = ypred;
          bool conv = false;
          for (int round = 0; round < 2 && !conv; ++round) {
            if (c != c_lu) {
              lin.factor(N, L, c, lane);
              c_lu = c;
            }
            double dy_old = -1.0, rate = -1.0;
            for (int it = 0; it < 4; ++it) {
              publish_y(N, L, ynew, lane);
              const double f = rhs_pieces(N, p, L, rc, lin.pw0, lin.pw1, lane);
              double dy = act ? c * f - psi - d : 0.0;
              dy = lin.solve(N, L, dy, lane);
              if (!act) dy = 0.0;
              const double dyn = sqrt(wave_sum(act ? (dy / scale) * (dy / scale) : 0.0) / N.ns);
              if (dy_old >= 0.0) rate = dyn / dy_old;
              // the step was accepted by the forward: a correction far below
              // the Newton tolerance ends the iteration even when rounding
              // makes the rate estimate meaningless
              const bool tiny = dyn < 1e-3 * newton_tol;
              if (rate >= 1.0 && !tiny) break;
              ynew += dy;
              d += dy;
              if (dyn == 0.0 || tiny ||
                  (rate >= 0.0 && rate / (1.0 - rate) * dyn < newton_tol)) {
                conv = true;
                break;
              }
              dy_old = dyn;
            }
            if (!conv) {  // refresh the Jacobian at the current iterate
              publish_y(N, L, ynew, lane);
              jacobian_sched<64>(N, p, L, lane);
              c_lu = -1.0;
            }
          }
          // the difference update of the step (as bdf_integrate)
#pragma unroll
          for (int j = kNumD - 1; j >= 2; --j)
            if (j == k + 2) D[j] = d - D[j - 1];
#pragma unroll
          for (int j = kNumD - 1; j >= 1; --j)
            if (j == k + 1) D[j] = d;
#pragma unroll
          for (int j = kNumD - 2; j >= 0; --j)
            if (j <= k) D[j] += D[j + 1];
          // factors of I - c J(y_n+1): kept in L.M for the next step's
          // Newton, stored in the column layout for the backward pass
          publish_y(N, L, D[0], lane);
          jacobian_sched<64>(N, p, L, lane);
          lin.factor(N, L, c, lane);
          c_lu = c;
          double *sp = seg + (size_t)(i - i0) * kTapeCol * 64;
#pragma unroll
          for (int q = 0; q < wave::kNS; ++q)
            sp[q * 64 + lane] = row ? L.M[wave::kPerm[q] * L.ld + lin.pos] : 0.0;
          sp[31 * 64 + lane] = row ? L.M[lane * L.ld + lin.pos] : 0.0;
          sp[32 * 64 + lane] = D[0];
        }
        // the recomputed segment must land on the forward's state (the next
        // checkpoint, or the integration's output after the last segment)
        // within 1e-6 relative + 1e-14 absolute (species far below the
        // abundances that matter differ at the atol level, where the floor
        // of the rate equations leaves the Newton iteration no grip)
        {
          const bool last = i1 >= nst;
          const double yref = last ? xe : (act ? ck[((sg + 1) * kNumD) * 32 + lane] : 0.0);
          const double ycur = last ? fmax(D[0], 1e-30) : D[0];
          const double dev = act ? fabs(ycur - yref) / (1e-6 * fabs(yref) + 1e-14) : 0.0;
          if (wave_max(dev) > 1.0 && failed == 0) failed = -3;
        }
        // backwards through the segment
        for (int i = i1 - 1; i >= i0; --i) {
          const double c = rec[3 * i], F = rec[3 * i + 1];
          const int k = (int)rec[3 * i + 2];
          // adjoint of the difference update: D'_k+1 = d, D'_k+2 = d - D_k+1,
          // D'_j = D_j + ... + D_k + d (j <= k)
          double db = 0.0, cum = 0.0, Dn[kNumD];
#pragma unroll
          for (int j = 0; j < kNumD; ++j) {
            if (j <= k) {
              cum += Db[j];
              Dn[j] = cum;
            } else if (j == k + 1) {
              Dn[j] = j + 1 < kNumD ? -Db[j + 1 < kNumD ? j + 1 : j] : 0.0;
            } else if (j == k + 2) {
              Dn[j] = 0.0;
            } else {
              Dn[j] = Db[j];
            }
          }
#pragma unroll
          for (int j = 0; j < kNumD; ++j)
            if (j <= k + 2) db += Db[j];
          // lambda = (I - c J)^-T d'
          const double *sp = seg + (size_t)(i - i0) * kTapeCol * 64;
          double at[32];
#pragma unroll
          for (int q = 0; q < 32; ++q) at[q] = (q < wave::kNS) ? sp[q * 64 + lane] : 0.0;
          const double dinv = sp[31 * 64 + lane];
          const double ynew = sp[32 * 64 + lane];
          double lam = wave::lu_solve_t(at, tlm, tum, lane, act ? db : 0.0, WaveBcast{}, L.xs,
                                        lin.pos, tlev, dinv);
          if (!act) lam = 0.0;
#pragma unroll
          for (int j = 0; j < kNumD; ++j) {
            if (j <= k) Dn[j] += lam - db;
            if (j >= 1 && j <= k) Dn[j] -= gam(j) / alp(k) * lam;
          }
          // parameters: c (df/dtheta)^T lambda at y_n+1
          publish_y(N, L, ynew, lane);
          L.psum[lane] = lam;
          __syncthreads();
          param_adjoint(N, p, L, lane, c, gk, sG, sZ, sN);
#pragma unroll
          for (int j = 0; j < kNumD; ++j) Db[j] = Dn[j];
          if (F != 1.0) change_D_adj(L, lane, Db, k, F);
        }
      }
      // D[1] = h0 f(y0)
      const double yi = act ? y0[(int64_t)lane * n + cell] : 0.0;
      publish_y(N, L, yi, lane);
      L.psum[lane] = act ? Db[1] : 0.0;
      __syncthreads();
      param_adjoint(N, p, L, lane, TD.hinit[cell], gk, sG, sZ, sN);
    }
    sG = wave_sum(sG);
    sZ = wave_sum(sZ);
    sN = wave_sum(sN);
    double *gout = grad + cell * (int64_t)(N.nr + 3);
#pragma unroll
    for (int q = 0; q < kRxPerLane; ++q)
      if (lane + 64 * q < N.nr) gout[lane + 64 * q] = gk[q];
    if (lane == 0) {
      gout[N.nr] = sG;
      gout[N.nr + 1] = sZ;
      gout[N.nr + 2] = ob.explicit_n + sN;
      obj[2 * cell] = ob.lam;
      obj[2 * cell + 1] = ob.co;
      if (status_out) status_out[cell] = failed;
    }
    __syncthreads();
  }
}

// Persistent one-cell-per-lane solve (variant 4): sizes the resident grid,
// grows the handle's slot memory and launches.
static int lane_solve(Network *net, int64_t n, double *y, const double *n_h, const double *t_gas,
                      const double *g0, const double *a_v, const double *zeta,
                      const double *fsh_h2, const double *fsh_co, const p3d_solve_opts &o,
                      int32_t *converged, int32_t *stats, void *stream) {
  int dev = 0, cus = 0;
  P3D_CHECK_HIP(hipGetDevice(&dev));
  P3D_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int wpc = o.waves_per_cu > 0 ? o.waves_per_cu : 8;
  int64_t n_waves = std::min<int64_t>((n + 63) / 64, (int64_t)cus * wpc);
  // slot memory: one 64-lane block per resident wavefront (bounded to 16 GiB)
  const int64_t max_waves = ((int64_t)16 << 30) / (64 * (int64_t)lane_slot_doubles() * 8);
  n_waves = std::max<int64_t>(1, std::min(n_waves, max_waves));
  const int64_t n_lanes = n_waves * 64;
  const size_t slot_bytes = (size_t)n_lanes * lane_slot_doubles() * sizeof(double);
  // slot memory, then the counters (64-bit queue head, 32-bit exited lanes,
  // 32-bit rescue count), then the rescue list (one int64 per lane)
  const size_t need = slot_bytes + 256 + (size_t)n_lanes * sizeof(int64_t);
  if (net->lane_ws_bytes < need) {
    if (net->lane_ws) {
      P3D_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
      P3D_CHECK_HIP(hipFree(net->lane_ws));
      net->lane_ws = nullptr;
      net->lane_ws_bytes = 0;
    }
    if (hipMalloc(&net->lane_ws, need) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipMalloc of %zu bytes of lane slot memory failed", need);
      return P3D_ENOMEM;
    }
    net->lane_ws_bytes = need;
  }
  LaneArgs a;
  a.n = n; a.y = y; a.nH = n_h; a.T = t_gas; a.G0 = g0; a.AV = a_v; a.zeta = zeta;
  a.fH2 = fsh_h2; a.fCO = fsh_co; a.conv = converged; a.stats = stats;
  a.rtype = net->d.rtype; a.alpha = net->d.alpha; a.beta = net->d.beta; a.gamma = net->d.gamma;
  a.post_w = net->d.post_w; a.opts = o; a.stream = stream;
  double *ws = static_cast<double *>(net->lane_ws);
  char *tail = static_cast<char *>(net->lane_ws) + slot_bytes;
  LaneQueue q;
  q.head = reinterpret_cast<unsigned long long *>(tail);
  q.exited = reinterpret_cast<unsigned *>(tail + 8);
  q.n_rescue = reinterpret_cast<unsigned *>(tail + 12);
  q.rescue = reinterpret_cast<int64_t *>(tail + 256);
  // once this many lanes found the queue empty, cells still integrating are
  // handed to the one-cell-per-workgroup kernel (long-tail cells)
  const double frac = o.rescue_frac > 0.0 ? std::min(o.rescue_frac, 1.0) : kLaneRescueFrac;
  q.exit_threshold = (unsigned)std::max<int64_t>(1, (int64_t)std::ceil(frac * (double)n_lanes));
  // wavefronts per SIMD the kernel is compiled for: 4 SIMDs per CU
  const int occ = std::min(4, std::max(2, (wpc + 3) / 4));
  int rc = lane_launch(a, n_waves, occ, ws, q);
  if (rc != P3D_OK) return rc;
  const int64_t max_rescue = n_lanes - (int64_t)q.exit_threshold + 1;
  const size_t sh = lds_bytes(net->d.ns, net->d.nr, net->d.n_rhs_e);
  P3D_REQUIRE(sh <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", sh);
  hipLaunchKernelGGL(chem_solve_kernel<LinReg<32>>, dim3((unsigned)max_rescue), dim3(64), sh,
                     as_stream(stream), net->d, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, o,
                     converged, stats, q.rescue, q.n_rescue);
  return launch_status("chem_solve_kernel (lane rescue)");
}

// Persistent one-cell-per-wavefront solve (variant 6): as many resident
// wavefronts as the occupancy allows (capped at waves_per_cu per CU when
// given), each taking cells from the handle's work counter.
static int queue_solve(Network *net, int64_t n, double *y, const double *n_h,
                       const double *t_gas, const double *g0, const double *a_v,
                       const double *zeta, const double *fsh_h2, const double *fsh_co,
                       const p3d_solve_opts &o, int32_t *converged, int32_t *stats,
                       void *stream) {
  const size_t sh = lds_bytes_wave();
  P3D_REQUIRE(sh <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", sh);
  if (!net->queue_ctr) {
    if (hipMalloc(&net->queue_ctr, 256) != hipSuccess) {
      (void)hipGetLastError();
      net->queue_ctr = nullptr;
      set_error("hipMalloc of the work counter failed");
      return P3D_ENOMEM;
    }
  }
  int dev = 0, cus = 0, per_cu = 0;
  P3D_CHECK_HIP(hipGetDevice(&dev));
  P3D_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  P3D_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(chem_queue_kernel<LinStatic>), 64, sh));
  if (o.waves_per_cu > 0) per_cu = std::min(per_cu, (int)o.waves_per_cu);
  per_cu = std::max(per_cu, 1);
  const int64_t waves = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)cus * per_cu));
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, net->queue_ctr);
  hipLaunchKernelGGL(chem_queue_kernel<LinStatic>, dim3((unsigned)waves), dim3(64), sh, s,
                     net->d, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, o, converged,
                     stats, net->queue_ctr);
  return launch_status("chem_queue_kernel");
}

// Tape workspace of n cells at cap steps per cell: step records, checkpoints,
// initial step sizes, step counts (8-byte aligned sections).
static size_t tape_section(int64_t n, int cap, int k) {
  const size_t rec = (size_t)n * cap * 3 * 8;
  const size_t ck = (size_t)n * (cap / kTapeSeg + 1) * kNumD * 32 * 8;
  const size_t h = (size_t)n * 8;
  const size_t ns = ((size_t)n * 4 + 7) / 8 * 8;
  const size_t sz[4] = {rec, ck, h, ns};
  size_t off = 0;
  for (int i = 0; i < k; ++i) off += sz[i];
  return off;
}
static size_t tape_bytes_of(int64_t n, int cap) { return tape_section(n, cap, 4); }
static TapeDev tape_view(void *ws, int64_t n, int cap) {
  char *b = static_cast<char *>(ws);
  TapeDev t;
  t.rec = reinterpret_cast<double *>(b + tape_section(n, cap, 0));
  t.ck = reinterpret_cast<double *>(b + tape_section(n, cap, 1));
  t.hinit = reinterpret_cast<double *>(b + tape_section(n, cap, 2));
  t.nsteps = reinterpret_cast<int32_t *>(b + tape_section(n, cap, 3));
  t.cap = cap;
  return t;
}

static int tape_checks(Network *net, int64_t n, const p3d_solve_opts *o, const void *tape,
                       int64_t tape_bytes, int32_t cap, const char *what) {
  P3D_REQUIRE(net->lane_ok,
              "%s: the discrete adjoint runs on the bundled network's register LU only", what);
  P3D_REQUIRE(o, "%s: solve options are required", what);
  P3D_REQUIRE(o->n_checkpoints == 2 && o->early_stop == 0 && o->post_mode == -1,
              "%s: the taped map is one span without post-solve corrections "
              "(n_checkpoints 2, early_stop 0, post_mode -1)", what);
  P3D_REQUIRE(o->rate_fp32 == 0, "%s: the taped map evaluates the rates in FP64", what);
  P3D_REQUIRE(cap >= kTapeSeg && cap % kTapeSeg == 0 && cap <= (1 << 24),
              "%s: steps per cell must be a multiple of %d in [%d, 2^24] (got %d)", what,
              kTapeSeg, kTapeSeg, cap);
  P3D_REQUIRE(tape && tape_bytes >= 0 && (size_t)tape_bytes >= tape_bytes_of(n, cap),
              "%s: the tape needs %zu bytes for %lld cells at %d steps (got %lld)", what,
              tape_bytes_of(n, cap), (long long)n, cap, (long long)tape_bytes);
  return P3D_OK;
}

// Resident wavefronts of a persistent one-cell-per-wavefront kernel.
static int resident_waves(const void *kernel, size_t sh, int64_t n, int64_t *waves) {
  int dev = 0, cus = 0, per_cu = 0;
  P3D_CHECK_HIP(hipGetDevice(&dev));
  P3D_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  P3D_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, sh));
  per_cu = std::max(per_cu, 1);
  *waves = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)cus * per_cu));
  return P3D_OK;
}

static int ensure_queue_ctr(Network *net) {
  if (!net->queue_ctr) {
    if (hipMalloc(&net->queue_ctr, 256) != hipSuccess) {
      (void)hipGetLastError();
      net->queue_ctr = nullptr;
      set_error("hipMalloc of the work counter failed");
      return P3D_ENOMEM;
    }
  }
  return P3D_OK;
}

}  // namespace p3d

using namespace p3d;

extern "C" {

// Static elimination schedule of LinSparseG from the Jacobian structure
// (plus the diagonal).  Order: greedy minimum Markowitz product over the
// remaining rows and columns, lowest species on ties, with the fill of each
// elimination added as it goes.  Level of pivot p: one above the highest
// level of the pivots k < p with a filled entry at (p, k) or (k, p), 0 when
// none.
struct SparseLuSched {
  int nlev = 0;
  std::vector<int> pos, lev, lvoff, lvp, uoff, uj;
};

static void sparse_lu_schedule(int ns, const std::vector<int> &jrow, const std::vector<int> &jcol,
                               SparseLuSched &s) {
  std::vector<char> A((size_t)ns * ns, 0);  // species coordinates
  for (int i = 0; i < ns; ++i) A[(size_t)i * ns + i] = 1;
  for (size_t z = 0; z < jrow.size(); ++z) A[(size_t)jrow[z] * ns + jcol[z]] = 1;
  std::vector<char> left(ns, 1);
  std::vector<int> order;
  for (int step = 0; step < ns; ++step) {
    long best = -1;
    int bk = -1;
    for (int k = 0; k < ns; ++k) {
      if (!left[k]) continue;
      long rc = 0, cc = 0;
      for (int j = 0; j < ns; ++j)
        if (left[j] && j != k) {
          rc += A[(size_t)k * ns + j];
          cc += A[(size_t)j * ns + k];
        }
      if (best < 0 || rc * cc < best) {
        best = rc * cc;
        bk = k;
      }
    }
    order.push_back(bk);
    left[bk] = 0;
    for (int i = 0; i < ns; ++i)
      if (left[i] && A[(size_t)i * ns + bk])
        for (int j = 0; j < ns; ++j)
          if (left[j] && A[(size_t)bk * ns + j]) A[(size_t)i * ns + j] = 1;
  }
  s.pos.assign(ns, 0);
  for (int p = 0; p < ns; ++p) s.pos[order[p]] = p;
  auto F = [&](int p, int q) { return A[(size_t)order[p] * ns + order[q]] != 0; };
  s.lev.assign(ns, 0);
  for (int p = 0; p < ns; ++p)
    for (int k = 0; k < p; ++k)
      if (F(p, k) || F(k, p)) s.lev[p] = std::max(s.lev[p], s.lev[k] + 1);
  s.nlev = 0;
  for (int p = 0; p < ns; ++p) s.nlev = std::max(s.nlev, s.lev[p] + 1);
  s.lvoff.assign(1, 0);
  s.lvp.clear();
  for (int v = 0; v < s.nlev; ++v) {
    for (int p = 0; p < ns; ++p)
      if (s.lev[p] == v) s.lvp.push_back(p);
    s.lvoff.push_back((int)s.lvp.size());
  }
  s.uoff.assign(1, 0);
  s.uj.clear();
  for (int p = 0; p < ns; ++p) {
    for (int j = p + 1; j < ns; ++j)
      if (F(p, j)) s.uj.push_back(j);
    s.uoff.push_back((int)s.uj.size());
  }
}

// The schedule packed for LDS (LinSparseG::stage): descriptor q of the
// pivots in level order = k | len << 7 | ubeg << 14 | (last of its level)
// << 31; the len upper positions of pivot k as bytes from ub[ubeg] (each
// list starting on a 4-byte boundary); zb[z] = position of column jcol[z];
// pb[s] = position of species s (kMaxSpeciesWide bytes).  Byte arrays are
// padded to whole words.  False when a field would not fit its bits.
struct SparseLuPacked {
  std::vector<unsigned> desc;
  std::vector<unsigned char> ub, zb, pb;
};

static bool pack_sparse_lu(int ns, const std::vector<int> &jcol, const SparseLuSched &sl,
                           SparseLuPacked &pk) {
  pk.desc.assign(ns, 0u);
  pk.ub.clear();
  pk.zb.assign(jcol.size(), 0);
  pk.pb.assign(kMaxSpeciesWide, 0);
  for (int v = 0; v < sl.nlev; ++v)
    for (int q = sl.lvoff[v]; q < sl.lvoff[v + 1]; ++q) {
      const int k = sl.lvp[q], len = sl.uoff[k + 1] - sl.uoff[k];
      while (pk.ub.size() % 4) pk.ub.push_back(0);
      const size_t ubeg = pk.ub.size();
      for (int u = sl.uoff[k]; u < sl.uoff[k + 1]; ++u) pk.ub.push_back((unsigned char)sl.uj[u]);
      if (k >= 128 || len >= 128 || ubeg >= (1u << 17)) return false;
      pk.desc[q] = (unsigned)k | ((unsigned)len << 7) | ((unsigned)ubeg << 14) |
                   (q == sl.lvoff[v + 1] - 1 ? 1u << 31 : 0u);
    }
  do pk.ub.push_back(0); while (pk.ub.size() % 4);
  for (size_t z = 0; z < jcol.size(); ++z) pk.zb[z] = (unsigned char)sl.pos[jcol[z]];
  do pk.zb.push_back(0); while (pk.zb.size() % 4);
  for (int i = 0; i < ns; ++i) pk.pb[i] = (unsigned char)sl.pos[i];
  return true;
}

int p3d_network_create(int32_t ns, int32_t nr, const int32_t *rtype, const double *alpha,
                       const double *beta, const double *gamma, const int32_t *reactants,
                       const int32_t *products, int32_t e_index, const int32_t *roles7,
                       const double *post_w4, int64_t *handle) {
  P3D_REQUIRE(handle, "handle pointer is required");
  P3D_REQUIRE(ns >= 1 && ns <= kMaxSpeciesWide, "n_species must be in [1, %d] (got %d)",
              kMaxSpeciesWide, ns);
  P3D_REQUIRE(nr >= 1 && nr <= 16384, "n_reactions must be in [1, 16384] (got %d)", nr);
  P3D_REQUIRE(rtype && alpha && beta && gamma && reactants && products, "null table");
  P3D_REQUIRE(e_index >= 0 && e_index < ns, "e_index out of range");
  for (int r = 0; r < nr; ++r) {
    P3D_REQUIRE(rtype[r] >= P3D_RX_TWO_BODY && rtype[r] <= P3D_RX_GRAIN,
                "reaction %d: unknown type %d", r, rtype[r]);
    const int a = reactants[2 * r], b = reactants[2 * r + 1];
    P3D_REQUIRE(a >= 0 && a < ns, "reaction %d: first reactant must be a species", r);
    if (rtype[r] == P3D_RX_TWO_BODY)
      P3D_REQUIRE((b >= 0 && b < ns) || b == P3D_SP_PAH_MINUS,
                  "reaction %d: two-body reaction needs a second reactant", r);
    else
      P3D_REQUIRE(b == P3D_SP_NONE || (b >= 0 && b < ns) || b == P3D_SP_PAH_MINUS,
                  "reaction %d: bad second reactant", r);
    for (int m = 0; m < 3; ++m)
      P3D_REQUIRE(products[3 * r + m] >= -2 && products[3 * r + m] < ns,
                  "reaction %d: product index out of range", r);
  }
  // dependency species of each rate
  std::vector<int> dep(2 * nr, -1);
  for (int r = 0; r < nr; ++r) {
    const int a = reactants[2 * r], b = reactants[2 * r + 1];
    dep[2 * r] = a;
    if (rtype[r] == P3D_RX_TWO_BODY) {
      if (b == P3D_SP_PAH_MINUS) dep[2 * r + 1] = e_index;
      else if (b != a) dep[2 * r + 1] = b;
    }
  }
  // RHS CSR in the upstream accumulation order: reaction by reaction,
  // reactant occurrences (-1) then product occurrences (+1).
  std::vector<std::vector<std::pair<int, double>>> rows(ns);
  for (int r = 0; r < nr; ++r) {
    const int slots = rtype[r] == P3D_RX_TWO_BODY || rtype[r] == P3D_RX_GRAIN ? 2 : 1;
    for (int m = 0; m < slots; ++m) {
      int s = reactants[2 * r + m];
      if (s >= 0) rows[s].push_back({r, -1.0});
    }
    for (int m = 0; m < 3; ++m) {
      int s = products[3 * r + m];
      if (s >= 0) rows[s].push_back({r, +1.0});
    }
  }
  // the same occurrences per reaction: consumed (2 slots), produced (3)
  std::vector<int> st5(5 * (size_t)nr, -1);
  for (int r = 0; r < nr; ++r) {
    const int slots = rtype[r] == P3D_RX_TWO_BODY || rtype[r] == P3D_RX_GRAIN ? 2 : 1;
    for (int m = 0; m < slots; ++m)
      if (reactants[2 * r + m] >= 0) st5[5 * r + m] = reactants[2 * r + m];
    for (int m = 0; m < 3; ++m)
      if (products[3 * r + m] >= 0) st5[5 * r + 2 + m] = products[3 * r + m];
  }
  std::vector<int> rhs_off(ns + 1, 0), rhs_rxn, rhs_pk;
  std::vector<double> rhs_sgn;
  for (int s = 0; s < ns; ++s) {
    for (auto &e : rows[s]) {
      rhs_rxn.push_back(e.first);
      rhs_sgn.push_back(e.second);
      rhs_pk.push_back((e.first << 1) | (e.second > 0.0 ? 1 : 0));
    }
    rhs_off[s + 1] = (int)rhs_rxn.size();
  }
  // Jacobian structure: entry (i, j) collects stoich(i, r) * d rate_r / d y_j
  std::map<std::pair<int, int>, std::vector<std::pair<int, double>>> jm;  // -> (2r+slot, coef)
  for (int r = 0; r < nr; ++r) {
    std::map<int, double> net;
    const int slots = rtype[r] == P3D_RX_TWO_BODY || rtype[r] == P3D_RX_GRAIN ? 2 : 1;
    for (int m = 0; m < slots; ++m)
      if (reactants[2 * r + m] >= 0) net[reactants[2 * r + m]] -= 1.0;
    for (int m = 0; m < 3; ++m)
      if (products[3 * r + m] >= 0) net[products[3 * r + m]] += 1.0;
    for (int sl = 0; sl < 2; ++sl) {
      const int j = dep[2 * r + sl];
      if (j < 0) continue;
      for (auto &kv : net)
        if (kv.second != 0.0) jm[{kv.first, j}].push_back({2 * r + sl, kv.second});
    }
  }
  std::vector<int> jrow, jcol, joff(1, 0), jt_rxn, jt_slot;
  std::vector<double> jt_coef;
  for (auto &kv : jm) {
    jrow.push_back(kv.first.first);
    jcol.push_back(kv.first.second);
    for (auto &t : kv.second) {
      jt_rxn.push_back(t.first / 2);
      jt_slot.push_back(t.first % 2);
      jt_coef.push_back(t.second);
    }
    joff.push_back((int)jt_rxn.size());
  }
  const int nnz = (int)jrow.size();
  std::vector<int> jr_start(ns + 1, nnz);
  std::vector<unsigned> jr_mask(ns, 0u);
  for (int z = nnz - 1; z >= 0; --z) jr_start[jrow[z]] = z;
  for (int r = ns - 1; r >= 0; --r)
    if (jr_start[r] > jr_start[r + 1]) jr_start[r] = jr_start[r + 1];
  for (int z = 0; z < nnz; ++z)
    if (jcol[z] < 32) jr_mask[jrow[z]] |= 1u << jcol[z];
  // balanced Jacobian schedule (longest-processing-time assignment of whole
  // entries to the 64 lanes; ties to the lowest lane)
  bool sched_ok = nnz < 512;
  for (size_t t = 0; t < jt_rxn.size() && sched_ok; ++t)
    sched_ok = 2 * jt_rxn[t] + jt_slot[t] < 256 && jt_coef[t] == std::nearbyint(jt_coef[t]) &&
               std::fabs(jt_coef[t]) <= 127.0;
  int jslots = 0;
  std::vector<unsigned> jsched;
  if (sched_ok) {
    std::vector<int> order(nnz);
    for (int z = 0; z < nnz; ++z) order[z] = z;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return joff[a + 1] - joff[a] > joff[b + 1] - joff[b];
    });
    std::vector<int> load(64, 0);
    std::vector<std::vector<int>> mine(64);
    for (int z : order) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      mine[l].push_back(z);
      load[l] += joff[z + 1] - joff[z];
    }
    jslots = *std::max_element(load.begin(), load.end());
    jsched.assign(64 * (size_t)std::max(jslots, 1), 0u);
    for (int l = 0; l < 64; ++l) {
      int sl = 0;
      for (int z : mine[l])
        for (int t = joff[z]; t < joff[z + 1]; ++t, ++sl)
          jsched[64 * (size_t)sl + l] =
              (unsigned)(2 * jt_rxn[t] + jt_slot[t]) |
              ((unsigned)(uint8_t)(int8_t)jt_coef[t] << 8) | ((unsigned)z << 16) |
              ((t == joff[z + 1] - 1 ? 1u : 0u) << 25) | (1u << 26);
    }
  } else {
    jsched.assign(64, 0u);
  }
  // pieces of the balanced RHS gather: cuts at every species row start and
  // at 128 equal shares of the entries
  const int n_e = (int)rhs_pk.size();
  std::vector<int> lane_cut(kMaxSpeciesWide + 1);
  for (int l = 0; l <= kMaxSpeciesWide; ++l)
    lane_cut[l] = (int)((int64_t)n_e * l / kMaxSpeciesWide);
  std::vector<int> pc_beg(rhs_off.begin(), rhs_off.end() - 1);
  pc_beg.insert(pc_beg.end(), lane_cut.begin(), lane_cut.end() - 1);
  std::sort(pc_beg.begin(), pc_beg.end());
  pc_beg.erase(std::unique(pc_beg.begin(), pc_beg.end()), pc_beg.end());
  while (!pc_beg.empty() && pc_beg.back() >= n_e) pc_beg.pop_back();
  const int n_pc = (int)pc_beg.size();
  pc_beg.push_back(n_e);
  auto first_piece = [&](int e) {
    return (int)(std::lower_bound(pc_beg.begin(), pc_beg.end() - 1, e) - pc_beg.begin());
  };
  std::vector<int> pc_lane(kMaxSpeciesWide + 1), pc_row(ns + 1);
  for (int l = 0; l <= kMaxSpeciesWide; ++l) pc_lane[l] = first_piece(lane_cut[l]);
  for (int r = 0; r <= ns; ++r) pc_row[r] = first_piece(rhs_off[r]);
  P3D_REQUIRE(n_pc <= kMaxPieces, "internal: %d RHS pieces", n_pc);
  std::vector<double> post(4 * (size_t)ns, 0.0);
  if (post_w4) post.assign(post_w4, post_w4 + 4 * (size_t)ns);
  int roles[7] = {-1, -1, -1, -1, -1, -1, -1};
  if (roles7)
    for (int k = 0; k < 7; ++k) {
      P3D_REQUIRE(roles7[k] >= -1 && roles7[k] < ns, "roles7[%d] out of range", k);
      roles[k] = roles7[k];
    }

  // pack all tables into one device allocation (8-byte aligned sections)
  std::vector<char> host;
  auto put = [&](const void *src, size_t bytes) {
    size_t off = host.size();
    size_t pad = (8 - off % 8) % 8;
    host.resize(off + pad + bytes);
    if (bytes) memcpy(host.data() + off + pad, src, bytes);
    return off + pad;
  };
  std::vector<int> rt(rtype, rtype + nr), re(reactants, reactants + 2 * nr);
  size_t o_rt = put(rt.data(), 4 * nr), o_a = put(alpha, 8 * nr), o_b = put(beta, 8 * nr),
         o_g = put(gamma, 8 * nr), o_re = put(re.data(), 8 * nr), o_dep = put(dep.data(), 8 * nr),
         o_ro = put(rhs_off.data(), 4 * (ns + 1)), o_rr = put(rhs_rxn.data(), 4 * rhs_rxn.size()),
         o_rs = put(rhs_sgn.data(), 8 * rhs_sgn.size()), o_pk = put(rhs_pk.data(), 4 * rhs_pk.size()),
         o_jr = put(jrow.data(), 4 * nnz),
         o_jc = put(jcol.data(), 4 * nnz), o_jo = put(joff.data(), 4 * (nnz + 1)),
         o_tr = put(jt_rxn.data(), 4 * jt_rxn.size()), o_ts = put(jt_slot.data(), 4 * jt_slot.size()),
         o_tc = put(jt_coef.data(), 8 * jt_coef.size()), o_pw = put(post.data(), 8 * post.size()),
         o_pb = put(pc_beg.data(), 4 * pc_beg.size()), o_pl = put(pc_lane.data(), 4 * pc_lane.size()),
         o_pr = put(pc_row.data(), 4 * pc_row.size()), o_s5 = put(st5.data(), 4 * st5.size()),
         o_js = put(jr_start.data(), 4 * jr_start.size()),
         o_jm = put(jr_mask.data(), 4 * jr_mask.size()),
         o_jq = put(jsched.data(), 4 * jsched.size());
  SparseLuSched sl;
  sparse_lu_schedule(ns, jrow, jcol, sl);
  SparseLuPacked pk;
  P3D_REQUIRE(pack_sparse_lu(ns, jcol, sl, pk), "internal: sparse LU schedule too large");
  const std::vector<unsigned> &sdesc = pk.desc;
  const std::vector<unsigned char> &sub = pk.ub, &szb = pk.zb, &spb = pk.pb;
  const size_t o_sdd = put(sdesc.data(), 4 * sdesc.size()), o_sub = put(sub.data(), sub.size()),
               o_szb = put(szb.data(), szb.size()), o_spb = put(spb.data(), spb.size());
  size_t o_sp = put(sl.pos.data(), 4 * sl.pos.size()), o_sv = put(sl.lev.data(), 4 * sl.lev.size()),
         o_slo = put(sl.lvoff.data(), 4 * sl.lvoff.size()),
         o_slp = put(sl.lvp.data(), 4 * sl.lvp.size()), o_suo = put(sl.uoff.data(), 4 * sl.uoff.size()),
         o_suj = put(sl.uj.data(), 4 * sl.uj.size());
  auto net = std::make_unique<Network>();
  if (hipMalloc(&net->blob, host.size()) != hipSuccess) {
    set_error("hipMalloc of %zu bytes for network tables failed", host.size());
    return P3D_ENOMEM;
  }
  P3D_CHECK_HIP(hipMemcpy(net->blob, host.data(), host.size(), hipMemcpyHostToDevice));
  char *b = static_cast<char *>(net->blob);
  NetDev &d = net->d;
  d.ns = ns; d.nr = nr; d.e_idx = e_index; d.nnz = nnz; d.n_rhs_e = (int)rhs_pk.size();
  for (int k = 0; k < 7; ++k) d.roles[k] = roles[k];
  d.rtype = (const int *)(b + o_rt);
  d.alpha = (const double *)(b + o_a);
  d.beta = (const double *)(b + o_b);
  d.gamma = (const double *)(b + o_g);
  d.react = (const int *)(b + o_re);
  d.dep = (const int *)(b + o_dep);
  d.st5 = (const int *)(b + o_s5);
  d.jr_start = (const int *)(b + o_js);
  d.jr_mask = (const unsigned *)(b + o_jm);
  d.rhs_off = (const int *)(b + o_ro);
  d.rhs_rxn = (const int *)(b + o_rr);
  d.rhs_sgn = (const double *)(b + o_rs);
  d.rhs_pk = (const int *)(b + o_pk);
  d.jrow = (const int *)(b + o_jr);
  d.jcol = (const int *)(b + o_jc);
  d.joff = (const int *)(b + o_jo);
  d.jt_rxn = (const int *)(b + o_tr);
  d.jt_slot = (const int *)(b + o_ts);
  d.jt_coef = (const double *)(b + o_tc);
  d.jslots = jslots;
  d.jsched = (const unsigned *)(b + o_jq);
  d.post_w = (const double *)(b + o_pw);
  d.n_pc = n_pc;
  d.pc_beg = (const int *)(b + o_pb);
  d.pc_lane = (const int *)(b + o_pl);
  d.pc_row = (const int *)(b + o_pr);
  d.sl_nlev = sl.nlev;
  d.sl_pos = (const int *)(b + o_sp);
  d.sl_lev = (const int *)(b + o_sv);
  d.sl_lvoff = (const int *)(b + o_slo);
  d.sl_lvp = (const int *)(b + o_slp);
  d.sl_uoff = (const int *)(b + o_suo);
  d.sl_uj = (const int *)(b + o_suj);
  d.sl_nub = (int)sub.size();
  d.sl_nzb = (int)szb.size();
  d.sl_desc = (const unsigned *)(b + o_sdd);
  d.sl_ub = (const unsigned *)(b + o_sub);
  d.sl_zb = (const unsigned *)(b + o_szb);
  d.sl_pb = (const unsigned *)(b + o_spb);
  // the generated register-LU code and its LDS layout fix the species,
  // reactions and Jacobian structure
  net->lane_ok = roles7 != nullptr &&
                 lane_network_matches(ns, nr, rtype, reactants, products, e_index, roles) &&
                 ns == wave::kNS && nr == wave::kNR && nnz == wave::kNNZ && sched_ok;
  std::lock_guard<std::mutex> lk(g_net_mu);
  int64_t h = g_next_handle++;
  g_nets[h] = std::move(net);
  *handle = h;
  return P3D_OK;
}

int p3d_sparse_lu_schedule(int32_t ns, int32_t nnz, const int32_t *jrow, const int32_t *jcol,
                           int32_t cap, int32_t *pos, int32_t *lev, int32_t *lvoff, int32_t *lvp,
                           int32_t *uoff, int32_t *uj, int32_t *counts) {
  P3D_REQUIRE(ns >= 1 && ns <= kMaxSpeciesWide, "n_species must be in [1, %d] (got %d)",
              kMaxSpeciesWide, ns);
  P3D_REQUIRE(nnz >= 0 && (nnz == 0 || (jrow && jcol)), "null structure");
  P3D_REQUIRE(pos && lev && lvoff && lvp && uoff && uj && counts, "null output");
  std::vector<int> r(jrow, jrow + nnz), c(jcol, jcol + nnz);
  for (int z = 0; z < nnz; ++z)
    P3D_REQUIRE(r[z] >= 0 && r[z] < ns && c[z] >= 0 && c[z] < ns, "entry %d out of range", z);
  SparseLuSched sl;
  sparse_lu_schedule(ns, r, c, sl);
  P3D_REQUIRE((int64_t)sl.uj.size() <= cap, "cap %d below the %zu upper entries", cap,
              sl.uj.size());
  std::copy(sl.pos.begin(), sl.pos.end(), pos);
  std::copy(sl.lev.begin(), sl.lev.end(), lev);
  std::copy(sl.lvoff.begin(), sl.lvoff.end(), lvoff);
  std::copy(sl.lvp.begin(), sl.lvp.end(), lvp);
  std::copy(sl.uoff.begin(), sl.uoff.end(), uoff);
  std::copy(sl.uj.begin(), sl.uj.end(), uj);
  counts[0] = sl.nlev;
  counts[1] = (int32_t)sl.uj.size();
  return P3D_OK;
}

int p3d_sparse_lu_packed(int32_t ns, int32_t nnz, const int32_t *jrow, const int32_t *jcol,
                         int32_t cap, uint32_t *desc, uint8_t *ub, uint8_t *zb, uint8_t *pb,
                         int32_t *sizes) {
  P3D_REQUIRE(ns >= 1 && ns <= kMaxSpeciesWide, "n_species must be in [1, %d] (got %d)",
              kMaxSpeciesWide, ns);
  P3D_REQUIRE(nnz >= 0 && (nnz == 0 || (jrow && jcol)), "null structure");
  P3D_REQUIRE(desc && ub && zb && pb && sizes, "null output");
  std::vector<int> r(jrow, jrow + nnz), c(jcol, jcol + nnz);
  for (int z = 0; z < nnz; ++z)
    P3D_REQUIRE(r[z] >= 0 && r[z] < ns && c[z] >= 0 && c[z] < ns, "entry %d out of range", z);
  SparseLuSched sl;
  sparse_lu_schedule(ns, r, c, sl);
  SparseLuPacked pk;
  P3D_REQUIRE(pack_sparse_lu(ns, c, sl, pk), "sparse LU schedule too large to pack");
  P3D_REQUIRE((int64_t)pk.ub.size() <= cap && (int64_t)pk.zb.size() <= cap,
              "cap %d below the %zu / %zu packed bytes", cap, pk.ub.size(), pk.zb.size());
  std::copy(pk.desc.begin(), pk.desc.end(), desc);
  std::copy(pk.ub.begin(), pk.ub.end(), ub);
  std::copy(pk.zb.begin(), pk.zb.end(), zb);
  std::copy(pk.pb.begin(), pk.pb.end(), pb);
  sizes[0] = (int32_t)pk.ub.size();
  sizes[1] = (int32_t)pk.zb.size();
  return P3D_OK;
}

int p3d_network_destroy(int64_t handle) {
  std::lock_guard<std::mutex> lk(g_net_mu);
  auto it = g_nets.find(handle);
  P3D_REQUIRE(it != g_nets.end(), "unknown network handle %lld", (long long)handle);
  g_nets.erase(it);
  return P3D_OK;
}

// Entry points whose kernels map one species to one lane of a wavefront.
static int check_narrow(const Network *net, const char *what) {
  P3D_REQUIRE(net->d.ns <= kMaxSpecies, "%s supports networks of at most %d species (got %d)",
              what, kMaxSpecies, net->d.ns);
  return P3D_OK;
}

static int check_cells(const Network *net, int64_t n, const void *y, const double *nH,
                       const double *T, const double *G0, const double *AV, const double *zeta,
                       const double *fH2, const double *fCO) {
  P3D_REQUIRE(net, "unknown network handle");
  P3D_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX * 64, "n out of range");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(y && nH && T && G0 && AV && zeta && fH2 && fCO, "null device array");
  return P3D_OK;
}

int p3d_chem_rhs(int64_t handle, int64_t n, const double *y, const double *n_h,
                 const double *t_gas, const double *g0, const double *a_v, const double *zeta,
                 const double *fsh_h2, const double *fsh_co, double *dydt, void *stream) {
  Network *net = find_net(handle);
  int rc = check_cells(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co);
  if (rc != P3D_OK || n == 0) return rc;
  if ((rc = check_narrow(net, "p3d_chem_rhs")) != P3D_OK) return rc;
  P3D_REQUIRE(dydt, "null output");
  const size_t sh = lds_bytes(net->d.ns, net->d.nr, net->d.n_rhs_e);
  P3D_REQUIRE(sh <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", sh);
  hipLaunchKernelGGL(chem_rhs_kernel, dim3((unsigned)n), dim3(64), sh, as_stream(stream), net->d,
                     n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, dydt);
  return launch_status("chem_rhs_kernel");
}

int p3d_chem_jacobian(int64_t handle, int64_t n, const double *y, const double *n_h,
                      const double *t_gas, const double *g0, const double *a_v,
                      const double *zeta, const double *fsh_h2, const double *fsh_co,
                      double *jac, void *stream) {
  Network *net = find_net(handle);
  int rc = check_cells(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co);
  if (rc != P3D_OK || n == 0) return rc;
  if ((rc = check_narrow(net, "p3d_chem_jacobian")) != P3D_OK) return rc;
  P3D_REQUIRE(jac, "null output");
  const size_t sh = lds_bytes(net->d.ns, net->d.nr, net->d.n_rhs_e);
  P3D_REQUIRE(sh <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", sh);
  hipLaunchKernelGGL(chem_jac_kernel, dim3((unsigned)n), dim3(64), sh, as_stream(stream), net->d,
                     n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, jac);
  return launch_status("chem_jac_kernel");
}

static int steady_cons_from(const p3d_steady_cons *c, const NetDev &d, SteadyCons &K) {
  P3D_REQUIRE(c, "constraint description (host) is required");
  P3D_REQUIRE(c->n_elem >= 0 && c->n_elem <= kMaxElem, "n_elem must be in [0, %d]", kMaxElem);
  K.n_elem = c->n_elem;
  for (int e = 0; e < kMaxElem; ++e)
    for (int s = 0; s < kMaxSpecies; ++s) K.nuclei[e][s] = e < c->n_elem ? c->nuclei[e][s] : 0;
  for (int s = 0; s < kMaxSpecies; ++s) K.cation[s] = s < d.ns ? c->cation[s] : 0;
  P3D_REQUIRE(d.e_idx >= 0 && d.e_idx < d.ns, "the network has no electron species");
  return P3D_OK;
}

int p3d_chem_polish(int64_t handle, int64_t n, double *y, const double *n_h, const double *t_gas,
                    const double *g0, const double *a_v, const double *zeta,
                    const double *fsh_h2, const double *fsh_co, const p3d_steady_cons *cons,
                    const double *y_totals, int32_t max_iter, int32_t *iters, void *stream) {
  Network *net = find_net(handle);
  int rc = check_cells(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co);
  if (rc != P3D_OK || n == 0) return rc;
  if ((rc = check_narrow(net, "p3d_chem_polish")) != P3D_OK) return rc;
  SteadyCons K;
  if ((rc = steady_cons_from(cons, net->d, K)) != P3D_OK) return rc;
  P3D_REQUIRE(max_iter >= 1 && max_iter <= 1000, "max_iter must be in [1, 1000]");
  const size_t sh = lds_bytes(net->d.ns, net->d.nr, net->d.n_rhs_e);
  P3D_REQUIRE(sh <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", sh);
  hipLaunchKernelGGL(chem_polish_kernel, dim3((unsigned)n), dim3(64), sh, as_stream(stream),
                     net->d, K, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, y_totals,
                     (int)max_iter, iters);
  return launch_status("chem_polish_kernel");
}

int p3d_chem_adjoint(int64_t handle, int64_t n, const double *y, const double *n_h,
                     const double *t_gas, const double *g0, const double *a_v,
                     const double *zeta, const double *fsh_h2, const double *fsh_co,
                     const p3d_steady_cons *cons, const double *cii5, double w_cii, double w_co,
                     double *obj, double *grad, void *stream) {
  Network *net = find_net(handle);
  int rc = check_cells(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co);
  if (rc != P3D_OK || n == 0) return rc;
  if ((rc = check_narrow(net, "p3d_chem_adjoint")) != P3D_OK) return rc;
  SteadyCons K;
  if ((rc = steady_cons_from(cons, net->d, K)) != P3D_OK) return rc;
  P3D_REQUIRE(cii5, "[CII] line parameters (host, 5 doubles) are required");
  P3D_REQUIRE(obj && grad, "null output");
  for (int k = 0; k < 7; ++k)
    P3D_REQUIRE(k == 2 || k == 4 || k == 6 || net->d.roles[k] >= 0,
                "the network lacks a species the objective needs (H, H2, C+, CO)");
  CiiLine cl;
  cl.freq = cii5[0]; cl.A = cii5[1]; cl.Tul = cii5[2]; cl.gu = cii5[3]; cl.gl = cii5[4];
  const size_t sh = lds_bytes(net->d.ns, net->d.nr, net->d.n_rhs_e);
  P3D_REQUIRE(sh <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", sh);
  hipLaunchKernelGGL(chem_adjoint_kernel, dim3((unsigned)n), dim3(64), sh, as_stream(stream),
                     net->d, K, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, cl, w_cii, w_co,
                     obj, grad);
  return launch_status("chem_adjoint_kernel");
}

int p3d_chem_tape_bytes(int64_t n, int32_t cap, int64_t *bytes) {
  P3D_REQUIRE(bytes, "bytes pointer is required");
  P3D_REQUIRE(n >= 0, "n must be non-negative");
  P3D_REQUIRE(cap >= kTapeSeg && cap % kTapeSeg == 0 && cap <= (1 << 24),
              "steps per cell must be a multiple of %d in [%d, 2^24] (got %d)", kTapeSeg,
              kTapeSeg, cap);
  *bytes = (int64_t)tape_bytes_of(n, cap);
  return P3D_OK;
}

int p3d_chem_solve_tape(int64_t handle, int64_t n, double *y, const double *n_h,
                        const double *t_gas, const double *g0, const double *a_v,
                        const double *zeta, const double *fsh_h2, const double *fsh_co,
                        const p3d_solve_opts *opts, void *tape, int64_t tape_bytes, int32_t cap,
                        int32_t *stats, void *stream) {
  Network *net = find_net(handle);
  int rc = check_cells(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co);
  if (rc != P3D_OK || n == 0) return rc;
  if ((rc = tape_checks(net, n, opts, tape, tape_bytes, cap, "p3d_chem_solve_tape")) != P3D_OK)
    return rc;
  const size_t sh = lds_bytes_wave();
  if ((rc = ensure_queue_ctr(net)) != P3D_OK) return rc;
  int64_t waves = 0;
  if ((rc = resident_waves(reinterpret_cast<const void *>(chem_tape_kernel), sh, n, &waves)) !=
      P3D_OK)
    return rc;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, net->queue_ctr);
  hipLaunchKernelGGL(chem_tape_kernel, dim3((unsigned)waves), dim3(64), sh, s, net->d, n, y, n_h,
                     t_gas, g0, a_v, zeta, fsh_h2, fsh_co, *opts, tape_view(tape, n, cap), stats,
                     net->queue_ctr);
  return launch_status("chem_tape_kernel");
}

int p3d_chem_tape_adjoint(int64_t handle, int64_t n, const double *y0, const double *y_end,
                          const double *n_h, const double *t_gas, const double *g0,
                          const double *a_v, const double *zeta, const double *fsh_h2,
                          const double *fsh_co, const p3d_solve_opts *opts, const void *tape,
                          int64_t tape_bytes, int32_t cap, const double *cii5, double w_cii,
                          double w_co, double *obj, double *grad, int32_t *status,
                          void *stream) {
  Network *net = find_net(handle);
  int rc = check_cells(net, n, y_end, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co);
  if (rc != P3D_OK || n == 0) return rc;
  if ((rc = tape_checks(net, n, opts, tape, tape_bytes, cap, "p3d_chem_tape_adjoint")) != P3D_OK)
    return rc;
  P3D_REQUIRE(y0, "the initial abundances y0 are required");
  P3D_REQUIRE(cii5, "[CII] line parameters (host, 5 doubles) are required");
  P3D_REQUIRE(obj && grad, "null output");
  for (int k = 0; k < 7; ++k)
    P3D_REQUIRE(k == 2 || k == 4 || k == 6 || net->d.roles[k] >= 0,
                "the network lacks a species the objective needs (H, H2, C+, CO)");
  CiiLine cl;
  cl.freq = cii5[0]; cl.A = cii5[1]; cl.Tul = cii5[2]; cl.gu = cii5[3]; cl.gl = cii5[4];
  const size_t sh = lds_bytes_wave();
  if ((rc = ensure_queue_ctr(net)) != P3D_OK) return rc;
  int64_t waves = 0;
  if ((rc = resident_waves(reinterpret_cast<const void *>(chem_tape_adjoint_kernel), sh, n,
                           &waves)) != P3D_OK)
    return rc;
  // per resident wavefront: the factors, 1/u and state of one segment's steps
  const size_t seg_bytes = (size_t)waves * kTapeSeg * kTapeCol * 64 * 8;
  if (net->seg_bytes < seg_bytes) {
    if (net->segbuf) (void)hipFree(net->segbuf);
    net->segbuf = nullptr;
    net->seg_bytes = 0;
    if (hipMalloc(&net->segbuf, seg_bytes) != hipSuccess) {
      (void)hipGetLastError();
      net->segbuf = nullptr;
      set_error("hipMalloc of %zu bytes for the adjoint's segment buffer failed", seg_bytes);
      return P3D_ENOMEM;
    }
    net->seg_bytes = seg_bytes;
  }
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(queue_reset, dim3(1), dim3(64), 0, s, net->queue_ctr);
  hipLaunchKernelGGL(chem_tape_adjoint_kernel, dim3((unsigned)waves), dim3(64), sh, s, net->d, n,
                     y0, y_end, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, *opts,
                     tape_view(const_cast<void *>(tape), n, cap), cl, w_cii, w_co, obj, grad,
                     status, static_cast<double *>(net->segbuf), net->queue_ctr);
  return launch_status("chem_tape_adjoint_kernel");
}

int p3d_chem_post(int64_t handle, int64_t n, double *y, const double *n_h,
                  const double *t_gas, const double *zeta, int32_t post_mode, double x_c_total,
                  double x_o_total, void *stream) {
  Network *net = find_net(handle);
  P3D_REQUIRE(net, "unknown network handle");
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  if (int rc = check_narrow(net, "p3d_chem_post")) return rc;
  P3D_REQUIRE(y && n_h && t_gas && zeta, "null device array");
  P3D_REQUIRE(post_mode == 0 || post_mode == 1, "post_mode must be 0 or 1");
  for (int k = 0; k < 7; ++k)
    P3D_REQUIRE(net->d.roles[k] >= 0, "the network has no species roles for the corrections");
  hipLaunchKernelGGL(chem_post_kernel, dim3((unsigned)n), dim3(64), 0, as_stream(stream), net->d,
                     n, y, n_h, t_gas, zeta, (int)post_mode, x_c_total, x_o_total);
  return launch_status("chem_post_kernel");
}

int p3d_chem_solve(int64_t handle, int64_t n, double *y, const double *n_h, const double *t_gas,
                   const double *g0, const double *a_v, const double *zeta,
                   const double *fsh_h2, const double *fsh_co, const p3d_solve_opts *opts,
                   int32_t *converged, int32_t *stats, void *stream) {
  Network *net = find_net(handle);
  int rc = check_cells(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co);
  if (rc != P3D_OK || n == 0) return rc;
  P3D_REQUIRE(opts, "opts is required");
  const p3d_solve_opts o = *opts;
  P3D_REQUIRE(o.n_checkpoints >= 2 && o.n_checkpoints <= 10000, "n_checkpoints must be >= 2");
  P3D_REQUIRE(o.t_start > 0.0, "t_start must be > 0");
  P3D_REQUIRE(o.t_max <= 0.0 || o.t_max > o.t_start, "t_max must exceed t_start");
  P3D_REQUIRE(o.rtol > 0.0 && o.atol > 0.0, "tolerances must be > 0");
  P3D_REQUIRE(o.max_steps > 0, "max_steps must be > 0");
  P3D_REQUIRE(o.post_mode >= -1 && o.post_mode <= 1, "post_mode must be -1, 0 or 1");
  if (o.post_mode >= 0 || o.early_stop)
    for (int k = 0; k < 7; ++k)
      P3D_REQUIRE(net->d.roles[k] >= 0,
                  "post-solve corrections / early stop need the species roles of the network");
  P3D_REQUIRE(o.variant >= 0 && o.variant <= 7, "variant must be 0 .. 7");
  P3D_REQUIRE(o.fast == 0 || o.fast == 1, "fast must be 0 or 1");
  if (o.fast) {
    P3D_REQUIRE(net->d.ns <= kMaxSpecies, "the fast mode takes at most %d species", kMaxSpecies);
    const size_t shf = lds_bytes(net->d.ns, net->d.nr, net->d.n_rhs_e);
    P3D_REQUIRE(shf <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", shf);
    hipLaunchKernelGGL(chem_fast_kernel, dim3((unsigned)n), dim3(64), shf, as_stream(stream),
                       net->d, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, o, converged,
                       stats);
    return launch_status("chem_fast_kernel");
  }
  P3D_REQUIRE((o.variant != 2 && o.variant != 3) || net->d.ns <= 32,
              "variants 2 and 3 need n_species <= 32");
  P3D_REQUIRE(o.variant != 4 || net->lane_ok,
              "variant 4 (one cell per lane) is generated for the bundled network structure; "
              "this network differs");
  P3D_REQUIRE(o.variant != 6 || net->lane_ok,
              "variant 6 (register LU, work queue) is generated for the bundled network "
              "structure; this network differs");
  P3D_REQUIRE(o.waves_per_cu >= 0 && o.waves_per_cu <= 64, "waves_per_cu must be in [0, 64]");
  P3D_REQUIRE(o.rescue_frac >= 0.0 && o.rescue_frac <= 1.0, "rescue_frac must be in [0, 1]");
  if (o.variant == 4) return lane_solve(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, o,
                                        converged, stats, stream);
  // auto: for the bundled network structure the persistent register-LU form
  if (o.variant == 6 || (o.variant == 0 && net->lane_ok))
    return queue_solve(net, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co, o, converged,
                       stats, stream);
  // auto from 64 species: the static sparse LU by levels (7), measured 1.2x
  // the MFMA panel LU (5) on the 64- and 103-species synthetic networks with
  // the same step counts
  if (net->d.ns > kMaxSpecies || o.variant == 7 || (o.variant == 0 && net->d.ns >= 64)) {
    // two wavefronts per cell, lane = species: the MFMA panel LU with
    // partial pivoting (5), or the static sparse LU by levels (7)
    P3D_REQUIRE(net->d.ns <= kMaxSpecies || o.variant == 0 || o.variant == 5 || o.variant == 7,
                "networks of more than %d species use the two-wavefront forms "
                "(variant 0, 5 or 7; got %d)", kMaxSpecies, o.variant);
    const bool sparse = o.variant == 7 || o.variant == 0;
    const size_t shw = sparse ? lds_bytes_wide_sparse(net->d)
                              : lds_bytes_wide(net->d.ns, net->d.nr, net->d.nnz);
    P3D_REQUIRE(shw <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", shw);
    if (sparse)
      hipLaunchKernelGGL((chem_solve_kernel<LinSparseG<Wave2>, Wave2>), dim3((unsigned)n),
                         dim3(Wave2::W), shw, as_stream(stream), net->d, n, y, n_h, t_gas, g0,
                         a_v, zeta, fsh_h2, fsh_co, o, converged, stats, nullptr, nullptr);
    else
      hipLaunchKernelGGL((chem_solve_kernel<LinMfmaG<Wave2, true>, Wave2>), dim3((unsigned)n),
                         dim3(Wave2::W), shw, as_stream(stream), net->d, n, y, n_h, t_gas, g0,
                         a_v, zeta, fsh_h2, fsh_co, o, converged, stats, nullptr, nullptr);
    return launch_status("chem_solve_kernel");
  }
  const size_t sh = lds_bytes(net->d.ns, net->d.nr, net->d.n_rhs_e);
  P3D_REQUIRE(sh <= 160 * 1024, "network too large for one LDS partition (%zu bytes)", sh);
  int var = o.variant;
  // auto below 64 species: the LU in LDS.  The register LU (2) gives
  // the same bits but needs 412 registers, one wavefront per SIMD: on the
  // 64^3 workload's converged inputs it took 10.1 s against 7.1 s for the
  // LDS LU (191 VGPRs, two per SIMD, no scratch).
  if (var == 0) var = 1;
  if (var == 5)
    hipLaunchKernelGGL(chem_solve_kernel<LinMfma>, dim3((unsigned)n), dim3(64), sh,
                       as_stream(stream), net->d, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co,
                       o, converged, stats, nullptr, nullptr);
  else if (var == 3)
    hipLaunchKernelGGL(chem_solve_kernel<LinLdsU<32>>, dim3((unsigned)n), dim3(64), sh,
                       as_stream(stream), net->d, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co,
                       o, converged, stats, nullptr, nullptr);
  else if (var == 2)
    hipLaunchKernelGGL(chem_solve_kernel<LinReg<32>>, dim3((unsigned)n), dim3(64), sh,
                       as_stream(stream), net->d, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co,
                       o, converged, stats, nullptr, nullptr);
  else
    hipLaunchKernelGGL(chem_solve_kernel<LinLds>, dim3((unsigned)n), dim3(64), sh,
                       as_stream(stream), net->d, n, y, n_h, t_gas, g0, a_v, zeta, fsh_h2, fsh_co,
                       o, converged, stats, nullptr, nullptr);
  return launch_status("chem_solve_kernel");
}

}  // extern "C"
