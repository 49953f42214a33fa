// Host build of the one-cell-per-lane solver (prism3d_amd/csrc/lane_bdf.h)
// for CPU tests: the same trip-stepped code the gfx950 kernel runs, driven
// one cell at a time with a single slot.  Compiled by tests/test_lane_cpu.py.
#include <vector>

#include "../prism3d_amd/csrc/lane_bdf.h"

using namespace p3d::lane;

namespace {
struct HostSrc {
  int64_t n, next;
  double *y;
  const double *nH, *T, *G0, *AV, *zeta, *fH2, *fCO;
  int *conv, *stats;
  int64_t fetch() { return next < n ? next++ : -1; }
  bool should_abandon() const { return false; }
  void abandon(int64_t) {}
  void load(int64_t c, Cell &p, double (&yv)[kNS]) {
    p.nH = nH[c]; p.T = T[c]; p.G0 = G0[c]; p.AV = AV[c];
    p.zeta = zeta[c]; p.fH2 = fH2[c]; p.fCO = fCO[c];
    for (int s = 0; s < kNS; ++s) yv[s] = y[(int64_t)s * n + c];
  }
  void store(int64_t c, const double (&yv)[kNS], int cv, const int (&st)[4]) {
    for (int s = 0; s < kNS; ++s) y[(int64_t)s * n + c] = yv[s];
    if (conv) conv[c] = cv;
    if (stats)
      for (int k = 0; k < 4; ++k) stats[4 * c + k] = st[k];
  }
};
}  // namespace

extern "C" {

int lane_ns(void) { return kNS; }
int lane_slot_size(void) { return kSlotSize; }

// y: (kNS, n) species-major, updated in place.  Returns the number of trips.
int64_t lane_solve_host(int64_t n, double *y, const double *nH, const double *T,
                        const double *G0, const double *AV, const double *zeta,
                        const double *fH2, const double *fCO, const int *rtype,
                        const double *alpha, const double *beta, const double *gamma,
                        const double *post_w, double t_max, double t_start, int nk,
                        int early_stop, double rtol, double atol, double stop_rtol,
                        int post_mode, int max_steps, double x_c_total, double x_o_total,
                        int *conv, int *stats) {
  HostSrc src{n, 0, y, nH, T, G0, AV, zeta, fH2, fCO, conv, stats};
  Tables tb{rtype, alpha, beta, gamma, post_w};
  Opts o;
  o.t_max = t_max; o.t_start = t_start; o.rtol = rtol; o.atol = atol; o.stop_rtol = stop_rtol;
  o.x_c_total = x_c_total; o.x_o_total = x_o_total;
  o.newton_tol = fmax(10.0 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));
  o.nk = nk; o.early_stop = early_stop; o.post_mode = post_mode; o.max_steps = max_steps;
  o.rate_fp32 = 0;
  std::vector<double> mem(kSlotSize, 0.0);
  Slot m{mem.data(), 1};
  Lane L;
  lane_init(L);
  int64_t trips = 0;
  while (L.phase != kExit) {
    trip(L, m, tb, o, src);
    ++trips;
  }
  return trips;
}

// Building blocks for unit tests (one cell, slot stride 1).
void lane_rhs_host(const double *cell7, const double *k, const double *y, double *f) {
  std::vector<double> mem(kSlotSize, 0.0);
  Slot m{mem.data(), 1};
  for (int r = 0; r < kNR; ++r) m.k(r) = k[r];
  Cell p{cell7[0], cell7[1], cell7[2], cell7[3], cell7[4], cell7[5], cell7[6]};
  double fv[kNS];
  for (int s = 0; s < kNS; ++s) m.yn(s) = y[s];
  rhs(p, m, kOffYN, fv);
  for (int s = 0; s < kNS; ++s) f[s] = fv[s];
}

// The generated sparse Jacobian entries (entry order of the generator).
void lane_jac_host(const double *cell7, const double *k, const double *y, double *jz) {
  std::vector<double> mem(kSlotSize, 0.0);
  Slot m{mem.data(), 1};
  for (int r = 0; r < kNR; ++r) m.k(r) = k[r];
  Cell p{cell7[0], cell7[1], cell7[2], cell7[3], cell7[4], cell7[5], cell7[6]};
  for (int s = 0; s < kNS; ++s) m.yp(s) = y[s];
  jacobian(p, m, kOffYP);
  for (int z = 0; z < kNJ; ++z) jz[z] = m.J(z);
}

// Solve (I - c J) x = b with the generated no-pivot LU; J given as the
// generated sparse entries.
void lane_lu_solve_host(const double *jz, double c, double *b) {
  std::vector<double> mem(kSlotSize, 0.0);
  Slot m{mem.data(), 1};
  for (int z = 0; z < kNJ; ++z) m.J(z) = jz[z];
  lu_factor(m, c);
  double bv[kNS];
  for (int s = 0; s < kNS; ++s) bv[s] = b[s];
  lu_solve(m, bv);
  for (int s = 0; s < kNS; ++s) b[s] = bv[s];
}
}
