// One cell per lane: the per-cell steady-state chemistry solve as a state
// machine that advances one "trip" at a time.
//
// Every lane of a wavefront works on its own cell and keeps only scalars in
// registers; the per-cell arrays (rate coefficients, Jacobian, LU factors,
// backward differences, predictor, Newton iterate) live in a slot of device
// memory laid out entry-major (entry e of slot s at base[e * n_slots + s]),
// so a wavefront's 64 lanes touch 512 contiguous bytes per entry.
//
// A trip runs the blocks below in a fixed order; each lane executes only the
// blocks of its current phase, so lanes at different points of their solve
// share the wavefront (the blocks are masked, not serialised per lane).  A
// lane whose cell is finished fetches the next cell from the work source in
// the same trip.  The same code compiles for the host (tests run it on the
// CPU against the oracle) and for gfx950.
//
// Algorithm (paths relative to the upstream repository):
//   steady-state driver   prism3d/chemistry/solver.py:145-225 (log-spaced
//                         checkpoints from t_start to t_max, early stop on six
//                         key species after four spans, failed spans keep the
//                         state)
//   integrator            scipy solve_ivp(method='BDF') as called at
//                         solver.py:195-201: the NDF of Shampine & Reichelt
//                         (1997), orders 1-5, Newton with a stale Jacobian,
//                         new LU on every change of h / alpha
//   H+/e- correction      solver.py:227-281; conservation solver.py:345-388
// The linear solves use the statically ordered sparse LU of lane_net.h.
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef P3D_HD
#define P3D_HD
#endif
#ifndef P3D_INLINE
#define P3D_INLINE inline
#endif
// Compiler-only memory fence: keeps the straight-line network code from
// hoisting every slot load (or forwarding every stored value) across rows,
// which would exhaust the register file.
#define P3D_FENCE() __asm__ __volatile__("" ::: "memory")
// true if the predicate holds on any lane of the wavefront (one lane on the host)
#ifndef P3D_ANY
#define P3D_ANY(x) (x)
#endif

namespace p3d {
namespace lane {

struct Cell {
  double nH, T, G0, AV, zeta, fH2, fCO;
};

// PAH- abundance from the floored electron abundance (solver.py:74-87) and
// its derivative with respect to x_e.
P3D_HD P3D_INLINE void pah_minus(double ye_floored, const Cell &p, double &pm, double &dpm) {
  const double xpah = 1e-7;
  const double ye = fmax(ye_floored, 1e-10);
  const double ne = ye * p.nH;
  if (p.G0 > 0.0 && ne > 0.0) {
    const double psi = p.G0 * sqrt(p.T) / ne;
    const double fm = fmax(500.0 / (psi + 500.0) * 0.5, 0.0);
    pm = xpah * fm;
    dpm = ye_floored > 1e-10 ? xpah * 250.0 * psi / ((psi + 500.0) * (psi + 500.0) * ye) : 0.0;
  } else {
    pm = xpah * 0.4;
    dpm = 0.0;
  }
}

}  // namespace lane
}  // namespace p3d

#include "lane_net.h"

namespace p3d {
namespace lane {

constexpr int kMaxOrder = 5;
constexpr int kNumD = kMaxOrder + 3;
// species chunk sizes of the load-compute-store loops
constexpr int kChunk4 = 4, kChunk8 = 8;

// slot layout (entries)
constexpr int kOffK = 0;
constexpr int kOffJ = kOffK + kNR;
constexpr int kOffLU = kOffJ + kNJ;
constexpr int kOffD = kOffLU + kNLU;
constexpr int kOffYP = kOffD + kNumD * kNS;   // predictor
constexpr int kOffPSI = kOffYP + kNS;         // psi of the NDF formula
constexpr int kOffYC = kOffPSI + kNS;         // state at the last checkpoint
constexpr int kOffYN = kOffYC + kNS;          // Newton iterate
constexpr int kOffDD = kOffYN + kNS;          // accumulated Newton correction
constexpr int kSlotSize = kOffDD + kNS;

// Host slot: plain pointer arithmetic (the gfx950 kernel uses a buffer-
// resource slot with the same accessors, see lane_kernels.hip).
struct Slot {
  double *b;  // base + slot
  int64_t s;  // number of slots (entry stride)
  P3D_HD P3D_INLINE double &at(int e) const { return b[(int64_t)e * s]; }
  P3D_HD P3D_INLINE double &k(int q) const { return at(kOffK + q); }
  P3D_HD P3D_INLINE double &J(int z) const { return at(kOffJ + z); }
  P3D_HD P3D_INLINE double &LU(int z) const { return at(kOffLU + z); }
  P3D_HD P3D_INLINE double &D(int j, int i) const { return at(kOffD + j * kNS + i); }
  P3D_HD P3D_INLINE double &yp(int i) const { return at(kOffYP + i); }
  P3D_HD P3D_INLINE double &psi(int i) const { return at(kOffPSI + i); }
  P3D_HD P3D_INLINE double &yc(int i) const { return at(kOffYC + i); }
  P3D_HD P3D_INLINE double &yn(int i) const { return at(kOffYN + i); }
  P3D_HD P3D_INLINE double &dd(int i) const { return at(kOffDD + i); }
};

// Reaction tables of a network handle (device or host pointers).
struct Tables {
  const int *rtype;
  const double *alpha, *beta, *gamma;
  const double *post_w;  // 4 x kNS: H, C, O nuclei of the minor species, cations
};

struct Opts {
  double t_max, t_start, rtol, atol, stop_rtol, x_c_total, x_o_total, newton_tol;
  int nk, early_stop, post_mode, max_steps, rate_fp32;
};

// Rate coefficient k_r (solver.py:100-133; network.py:139-176).
P3D_HD P3D_INLINE double rate_coeff(int t, double a, double b, double g, const Cell &p,
                                    double Tc, double lTc, double lT) {
  switch (t) {
    case 0: return a * exp(b * lTc) * exp(-g / Tc);        // two-body
    case 1: return a * p.G0 * exp(-g * p.AV);               // photo
    case 2: return a * p.G0 * exp(-g * p.AV) * p.fH2;       // photo, H2 self-shielded
    case 3: return a * p.G0 * exp(-g * p.AV) * p.fCO;       // photo, CO shielded
    case 4: return a * p.zeta;                               // cosmic ray
    default: return a * exp(b * lT);                         // grain surface
  }
}

// The same with the exponentials in FP32 (one expf of the combined exponent).
P3D_HD P3D_INLINE double rate_coeff_f32(int t, double a, double b, double g, const Cell &p) {
  const float Tc = fminf(fmaxf((float)p.T, 10.0f), 41000.0f);
  const float bf = (float)b, gf = (float)g;
  switch (t) {
    case 0: return a * (double)expf(bf * logf(Tc / 300.0f) - gf / Tc);
    case 1: return a * p.G0 * (double)expf(-gf * (float)p.AV);
    case 2: return a * p.G0 * (double)expf(-gf * (float)p.AV) * p.fH2;
    case 3: return a * p.G0 * (double)expf(-gf * (float)p.AV) * p.fCO;
    case 4: return a * p.zeta;
    default: return a * (double)expf(bf * logf((float)p.T / 300.0f));
  }
}

// NDF coefficients (Shampine & Reichelt 1997): kappa_k, gamma_k = sum 1/j,
// alpha_k = (1 - kappa_k) gamma_k, error constant kappa_k gamma_k + 1/(k+1).
P3D_HD P3D_INLINE double ndf_gamma(int k) {
  return k == 1 ? 1.0 : k == 2 ? 1.5 : k == 3 ? 1.8333333333333333 : k == 4 ? 2.0833333333333333
         : k == 5 ? 2.2833333333333333 : k == 6 ? 2.45 : 0.0;
}
P3D_HD P3D_INLINE double ndf_kappa(int k) {
  return k == 1 ? -0.1850 : k == 2 ? -1.0 / 9.0 : k == 3 ? -0.0823 : k == 4 ? -0.0415 : 0.0;
}
P3D_HD P3D_INLINE double ndf_alpha(int k) { return (1.0 - ndf_kappa(k)) * ndf_gamma(k); }
P3D_HD P3D_INLINE double ndf_err(int k) { return ndf_kappa(k) * ndf_gamma(k) + 1.0 / (k + 1); }

// D <- (R(factor) U)^T D on rows 0..order of the differences:
// R_im = prod_{q=1..i} (q-1 - factor m)/q, U = R(1).
template <class S>
P3D_HD P3D_INLINE void change_D(const S &m, int order, double factor) {
  double RU[kMaxOrder + 1][kMaxOrder + 1];
#pragma unroll
  for (int i = 0; i <= kMaxOrder; ++i)
#pragma unroll
    for (int j = 0; j <= kMaxOrder; ++j) {
      double acc = 0.0;
      if (i <= order && j <= order) {
#pragma unroll
        for (int mm = 0; mm <= kMaxOrder; ++mm) {
          if (mm > order) break;
          double r = 1.0, u = 1.0;
#pragma unroll
          for (int q = 1; q <= kMaxOrder; ++q) {
            if (q <= i) r *= ((double)(q - 1) - factor * (double)mm) / (double)q;
            if (q <= mm) u *= ((double)(q - 1) - (double)j) / (double)q;
          }
          acc += r * u;
        }
      }
      RU[i][j] = acc;
    }
  // species in chunks: all loads of a chunk, then its stores (slot stores
  // would otherwise order every later load behind them)
  for (int c0 = 0; c0 < kNS; c0 += kChunk4) {
    double Dv[kChunk4][kMaxOrder + 1];
#pragma unroll
    for (int q = 0; q < kChunk4; ++q)
#pragma unroll
      for (int i = 0; i <= kMaxOrder; ++i)
        Dv[q][i] = (c0 + q < kNS && i <= order) ? (double)m.D(i, c0 + q) : 0.0;
#pragma unroll
    for (int q = 0; q < kChunk4; ++q) {
      if (c0 + q < kNS) {
#pragma unroll
        for (int j = 0; j <= kMaxOrder; ++j) {
          if (j <= order) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i <= kMaxOrder; ++i) acc += RU[i][j] * Dv[q][i];
            m.D(j, c0 + q) = acc;
          }
        }
      }
    }
  }
}

P3D_HD P3D_INLINE double rms_sum(double sumsq) { return sqrt(sumsq / kNS); }

// H+/e- correction (solver.py:227-281) and conservation (solver.py:345-388);
// mode 0 = upstream clipped scales, 1 = exact projection.
P3D_HD P3D_INLINE void post_solve(const Tables &tb, const Cell &p, double (&y)[kNS], int mode,
                                  double x_c_total, double x_o_total) {
  const int iH = kRoles[0], iH2 = kRoles[1], iHp = kRoles[2], iCp = kRoles[3], iC = kRoles[4],
            iCO = kRoles[5], iO = kRoles[6], ie = kE;
  const double *wH = tb.post_w, *wC = tb.post_w + kNS, *wO = tb.post_w + 2 * kNS,
               *wI = tb.post_w + 3 * kNS;
  {
    const double xH = y[iH], xCp = y[iCp];
    const double R = p.zeta * 0.46 * xH;
    const double Tc = fmax(p.T, 10.0);
    const double ae = 2.753e-14 * pow(Tc / 300.0, -0.745) + 1.0e-14 * sqrt(Tc / 100.0);
    const double a = ae * p.nH, b = ae * p.nH * xCp, c = -R;
    if (a > 0.0) {
      const double disc = b * b - 4.0 * a * c;
      if (disc >= 0.0 && xCp < 1e-6) {
        y[iHp] = fmax((-b + sqrt(disc)) / (2.0 * a), 1e-30);
        double npos = 0.0;
#pragma unroll
        for (int i = 0; i < kNS; ++i) npos += wI[i] * y[i];
        y[ie] = fmax(npos, 1e-30);
      }
    }
  }
  const double Hmaj = y[iH] + 2.0 * y[iH2];
  double Hmin = 0.0;
#pragma unroll
  for (int i = 0; i < kNS; ++i) Hmin += wH[i] * y[i];
  if (Hmaj > 0.0) {
    const double s = mode == 0 ? fmin(fmax(fmax(1.0 - Hmin, 0.9) / Hmaj, 0.8), 1.2)
                               : (1.0 - Hmin) / Hmaj;
    y[iH] *= s;
    y[iH2] *= s;
  }
  const double Cmaj = y[iCp] + y[iC] + y[iCO];
  double Cmin = 0.0;
#pragma unroll
  for (int i = 0; i < kNS; ++i) Cmin += wC[i] * y[i];
  if (Cmaj > 0.0) {
    const double s = mode == 0
                         ? fmin(fmax(fmax(x_c_total - Cmin, x_c_total * 0.9) / Cmaj, 0.8), 1.2)
                         : (x_c_total - Cmin) / Cmaj;
    y[iCp] *= s;
    y[iC] *= s;
    y[iCO] *= s;
  }
  double Omol = 0.0;
#pragma unroll
  for (int i = 0; i < kNS; ++i) Omol += wO[i] * y[i];
  y[iO] = fmax(x_o_total - Omol, 1e-30);
  double npos = 0.0;
#pragma unroll
  for (int i = 0; i < kNS; ++i) npos += wI[i] * y[i];
  y[ie] = fmax(npos, 1e-30);
#pragma unroll
  for (int i = 0; i < kNS; ++i) y[i] = fmax(y[i], 1e-30);
}

enum Phase : int {
  kFetch = 0,   // take the next cell from the source
  kSpan = 1,    // start the next checkpoint span
  kStep = 2,    // begin a step: clamp h
  kAttempt = 3, // rescale the differences, predict, new Newton solve
  kNewton = 4,  // one Newton iteration per trip
  kSpanEnd = 5, // span finished (or failed): checkpoint logic
  kFinish = 6,  // cell finished: post-solve corrections and output
  kExit = 7     // source exhausted
};

// Per-lane scalar state.
struct Lane {
  int64_t cell;
  int phase;
  Cell p;
  // checkpoint loop
  int span, conv, status, trivial;
  double tmax, l0, l1;
  // integration of one span
  double t, t1, h, max_step, hmin, tn, c, c_lu, dy_old, pend;
  int order, n_equal, it, steps_span, nres;
  int jac_current, lu_valid, need_lu, need_jac;
  int fresh;  // Newton iterate = predictor and correction = 0, not yet stored
  double err_sq;  // error-test sum of squares of the converged iterate
  // counters over the whole cell
  int st_steps, st_lu, st_rhs, st_jac;
};

P3D_HD P3D_INLINE void lane_init(Lane &L) {
  L.phase = kFetch;
  L.cell = -1;
}

// One trip of one lane.  Src provides:
//   int64_t fetch();                                       next cell or -1
//   bool should_abandon();    hand the current cell over (it is then solved
//   void abandon(int64_t);    elsewhere from its unchanged inputs)
//   void load(int64_t cell, Cell &p, double (&y)[kNS]);    inputs of a cell
//   void store(int64_t cell, const double (&y)[kNS], int conv, const int (&st)[4]);
template <class S, class Src>
P3D_HD P3D_INLINE void trip(Lane &L, const S &m, const Tables &tb, const Opts &o, Src &src) {
  // ---- hand a still-integrating cell over when the source asks for it -----
  if (L.phase >= kSpan && L.phase <= kSpanEnd && src.should_abandon()) {
    src.abandon(L.cell);
    L.phase = kExit;
  }
  // ---- finish a cell: corrections and output --------------------------
  P3D_FENCE();
  if (L.phase == kFinish) {
    double y[kNS];
#pragma unroll
    for (int i = 0; i < kNS; ++i) y[i] = m.yc(i);
    if (o.post_mode >= 0) post_solve(tb, L.p, y, o.post_mode, o.x_c_total, o.x_o_total);
    const int st[4] = {L.st_steps, L.st_lu, L.st_rhs, L.st_jac};
    src.store(L.cell, y, L.conv, st);
    L.phase = kFetch;
  }
  // ---- take a new cell ----------------------------------------------------
  P3D_FENCE();
  if (L.phase == kFetch) {
    const int64_t c = src.fetch();
    if (c < 0) {
      L.phase = kExit;
    } else {
      L.cell = c;
      double y[kNS];
      src.load(c, L.p, y);
#pragma unroll
      for (int i = 0; i < kNS; ++i) m.yc(i) = y[i];
      const double Tc = fmin(fmax(L.p.T, 10.0), 41000.0);
      const double lTc = log(Tc / 300.0), lT = log(L.p.T / 300.0);
      for (int r = 0; r < kNR; ++r)
        m.k(r) = o.rate_fp32
                     ? rate_coeff_f32(tb.rtype[r], tb.alpha[r], tb.beta[r], tb.gamma[r], L.p)
                     : rate_coeff(tb.rtype[r], tb.alpha[r], tb.beta[r], tb.gamma[r], L.p, Tc,
                                  lTc, lT);
      L.tmax = o.t_max > 0.0 ? o.t_max : fmin(3e17, fmax(1e16, 3e17 / fmax(L.p.nH, 1.0)));
      L.l0 = log10(o.t_start);
      L.l1 = log10(L.tmax);
      L.span = 0;
      L.conv = 0;
      L.st_steps = L.st_lu = L.st_rhs = L.st_jac = 0;
      L.phase = kSpan;
    }
  }
  // ---- start a span: initial step, differences, Jacobian ------------------
  P3D_FENCE();
  if (L.phase == kSpan) {
    if (L.span >= o.nk - 1) {
      L.phase = kFinish;
    } else {
      const int i = L.span, nk = o.nk;
      const double ta = i == 0 ? o.t_start : pow(10.0, L.l0 + (L.l1 - L.l0) * i / (nk - 1));
      const double tb_ = (i + 1 == nk - 1) ? L.tmax
                                           : pow(10.0, L.l0 + (L.l1 - L.l0) * (i + 1) / (nk - 1));
      const double span = tb_ - ta;
      L.status = 0;
      L.trivial = 0;
      if (!(span > 0.0)) {
        L.trivial = 1;
        L.phase = kSpanEnd;
      } else {
        L.t = ta;
        L.t1 = tb_;
        L.max_step = span;
        // f0 = f(y) with y = yc; f0 is kept in the psi entries (this block
        // runs once per span, so it trades memory traffic for registers)
        double h0, d1;
        {
          double f0[kNS], yv[kNS];
          rhs(L.p, m, kOffYC, f0);
#pragma unroll
          for (int s = 0; s < kNS; ++s) yv[s] = m.yc(s);
          // initial step (Hairer, Norsett & Wanner 1993, II.4), order 1
          double s0 = 0.0, s1 = 0.0;
#pragma unroll
          for (int s = 0; s < kNS; ++s) {
            const double sc = o.atol + o.rtol * fabs(yv[s]);
            s0 += (yv[s] / sc) * (yv[s] / sc);
            s1 += (f0[s] / sc) * (f0[s] / sc);
          }
          const double d0 = rms_sum(s0);
          d1 = rms_sum(s1);
          h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
          h0 = fmin(h0, span);
#pragma unroll
          for (int s = 0; s < kNS; ++s) {
            m.yn(s) = yv[s] + h0 * f0[s];
            m.psi(s) = f0[s];
          }
        }
        L.st_rhs++;
        P3D_FENCE();
        double s2 = 0.0;
        {
          double f1[kNS];
          rhs(L.p, m, kOffYN, f1);
          for (int c0 = 0; c0 < kNS; c0 += kChunk8) {
            double yv[kChunk8], fv[kChunk8];
#pragma unroll
            for (int u = 0; u < kChunk8; ++u)
              if (c0 + u < kNS) { yv[u] = m.yc(c0 + u); fv[u] = m.psi(c0 + u); }
#pragma unroll
            for (int u = 0; u < kChunk8; ++u)
              if (c0 + u < kNS) {
                const double sc = o.atol + o.rtol * fabs(yv[u]);
                double f1u = 0.0;
#pragma unroll
                for (int z = 0; z < kNS; ++z)
                  if (z == c0 + u) f1u = f1[z];
                const double df = f1u - fv[u];
                s2 += (df / sc) * (df / sc);
              }
          }
        }
        L.st_rhs++;
        const double d2 = rms_sum(s2) / h0;
        const double dm = fmax(d1, d2);
        const double h1 = (dm <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : sqrt(0.01 / dm);
        L.h = fmin(fmin(100.0 * h0, h1), fmin(L.max_step, span));
        for (int c0 = 0; c0 < kNS; c0 += kChunk8) {
          double yv[kChunk8], fv[kChunk8];
#pragma unroll
          for (int u = 0; u < kChunk8; ++u)
            if (c0 + u < kNS) { yv[u] = m.yc(c0 + u); fv[u] = m.psi(c0 + u); }
#pragma unroll
          for (int u = 0; u < kChunk8; ++u)
            if (c0 + u < kNS) {
              m.D(0, c0 + u) = yv[u];
              m.D(1, c0 + u) = fv[u] * L.h;
#pragma unroll
              for (int j = 2; j < kNumD; ++j) m.D(j, c0 + u) = 0.0;
            }
        }
        P3D_FENCE();
        L.order = 1;
        L.n_equal = 0;
        jacobian(L.p, m, kOffYC);
        L.st_jac++;
        L.jac_current = 1;
        L.lu_valid = 0;
        L.need_jac = 0;
        L.c_lu = 0.0;
        L.pend = 1.0;
        L.steps_span = 0;
        L.phase = kStep;
      }
    }
  }
  // ---- begin a step: clamp h to the span and the step limits --------------
  P3D_FENCE();
  if (L.phase == kStep) {
    if (!(L.t < L.t1)) {
      L.phase = kSpanEnd;
    } else if (L.steps_span++ > o.max_steps) {
      L.status = -2;
      L.phase = kSpanEnd;
    } else {
      L.hmin = 10.0 * fabs(nextafter(L.t, INFINITY) - L.t);
      if (L.h > L.max_step) {
        L.pend *= L.max_step / L.h; L.h = L.max_step; L.n_equal = 0; L.lu_valid = 0;
      }
      if (L.h < L.hmin) {
        L.pend *= L.hmin / L.h; L.h = L.hmin; L.n_equal = 0; L.lu_valid = 0;
      }
      if (L.t + L.h > L.t1) {
        L.pend *= (L.t1 - L.t) / L.h; L.h = L.t1 - L.t; L.n_equal = 0; L.lu_valid = 0;
      }
      L.phase = kAttempt;
    }
  }
  // ---- new attempt: rescale differences, predictor, Newton start ----------
  P3D_FENCE();
  if (L.phase == kAttempt) {
    if (L.pend != 1.0) {
      change_D(m, L.order, L.pend);
      L.pend = 1.0;
    }
    L.tn = fmin(L.t + L.h, L.t1);
    const double al = ndf_alpha(L.order);
    for (int c0 = 0; c0 < kNS; c0 += kChunk4) {
      double Dv[kChunk4][kMaxOrder + 1];
#pragma unroll
      for (int q = 0; q < kChunk4; ++q)
#pragma unroll
        for (int j = 0; j <= kMaxOrder; ++j)
          Dv[q][j] = (c0 + q < kNS && j <= L.order) ? (double)m.D(j, c0 + q) : 0.0;
#pragma unroll
      for (int q = 0; q < kChunk4; ++q) {
        if (c0 + q < kNS) {
          double yp = 0.0, ps = 0.0;
#pragma unroll
          for (int j = 0; j <= kMaxOrder; ++j)
            if (j <= L.order) yp += Dv[q][j];
#pragma unroll
          for (int j = 1; j <= kMaxOrder; ++j)
            if (j <= L.order) ps += ndf_gamma(j) * Dv[q][j];
          m.yp(c0 + q) = yp;
          m.psi(c0 + q) = ps / al;
        }
      }
    }
    L.c = L.h / ndf_alpha(L.order);
    L.need_lu = (!L.lu_valid || L.c != L.c_lu) ? 1 : 0;
    L.it = 0;
    L.dy_old = -1.0;
    L.fresh = 1;
    L.phase = kNewton;
  }
  // ---- Jacobian refresh at the predictor after a failed Newton solve ------
  P3D_FENCE();
  if (L.phase == kNewton && L.need_jac) {
    jacobian(L.p, m, kOffYP);
    L.fresh = 1;
    L.st_jac++;
    L.jac_current = 1;
    L.lu_valid = 0;
    L.need_jac = 0;
    L.need_lu = 1;
    L.it = 0;
    L.dy_old = -1.0;
  }
  // ---- LU of M = I - c J ---------------------------------------------------
  P3D_FENCE();
  if (L.phase == kNewton && L.need_lu) {
    lu_factor(m, L.c);
    L.st_lu++;
    L.lu_valid = 1;
    L.c_lu = L.c;
    L.need_lu = 0;
  }
  // ---- one Newton iteration ------------------------------------------------
  P3D_FENCE();
  if (L.phase == kNewton) L.nres = 0;
  // Newton iterations of the step: lanes iterate until converged (nres 1) or
  // failed (nres 2); the wavefront leaves the loop when no lane iterates, so
  // the predictor, factorisation and acceptance blocks run once per step
  for (int pass = 0; pass < 4 && P3D_ANY(L.phase == kNewton && L.nres == 0); ++pass) {
    P3D_FENCE();
    if (L.phase == kNewton && L.nres == 0) {
      double f[kNS];
      rhs(L.p, m, L.fresh ? kOffYP : kOffYN, f);
      L.st_rhs++;
      double fs = 0.0;
#pragma unroll
      for (int s = 0; s < kNS; ++s) fs += f[s];
      if (!isfinite(fs)) {
        L.nres = 2;
      } else {
#pragma unroll
        for (int s = 0; s < kNS; ++s) f[s] = L.c * f[s] - m.psi(s);
        if (!L.fresh) {
#pragma unroll
          for (int s = 0; s < kNS; ++s) f[s] -= m.dd(s);
        }
        lu_solve(m, f);
        double sq = 0.0;
#pragma unroll
        for (int s = 0; s < kNS; ++s) {
          const double sc = o.atol + o.rtol * fabs(m.yp(s));
          sq += (f[s] / sc) * (f[s] / sc);
        }
        const double dyn = rms_sum(sq);
        const double rate = L.dy_old >= 0.0 ? dyn / L.dy_old : -1.0;
        if (rate >= 0.0 &&
            (rate >= 1.0 || pow(rate, (double)(4 - L.it)) / (1.0 - rate) * dyn > o.newton_tol)) {
          L.nres = 2;
        } else {
          // converged with this update: the error test's sum is taken on the
          // updated iterate and correction while they are in registers
          const bool conv =
              dyn == 0.0 || (rate >= 0.0 && rate / (1.0 - rate) * dyn < o.newton_tol);
          const double ecq = ndf_err(L.order);
          double esq = 0.0;
#pragma unroll
          for (int c0 = 0; c0 < kNS; c0 += kChunk8) {
            double yv[kChunk8], dv[kChunk8];
            if (L.fresh) {
#pragma unroll
              for (int q = 0; q < kChunk8; ++q)
                if (c0 + q < kNS) { yv[q] = m.yp(c0 + q); dv[q] = 0.0; }
            } else {
#pragma unroll
              for (int q = 0; q < kChunk8; ++q)
                if (c0 + q < kNS) { yv[q] = m.yn(c0 + q); dv[q] = m.dd(c0 + q); }
            }
#pragma unroll
            for (int q = 0; q < kChunk8; ++q)
              if (c0 + q < kNS) {
                const double yn = yv[q] + f[c0 + q], dn = dv[q] + f[c0 + q];
                m.yn(c0 + q) = yn;
                m.dd(c0 + q) = dn;
                if (conv) {
                  const double sc = o.atol + o.rtol * fabs(yn);
                  const double e = ecq * dn;
                  esq += (e / sc) * (e / sc);
                }
              }
          }
          L.fresh = 0;
          if (conv) {
            L.err_sq = esq;
            L.nres = 1;
          } else {
            L.dy_old = dyn;
            L.it++;
            if (L.it == 4) L.nres = 2;
          }
        }
      }
    }
  }
  // ---- after the iteration: failure handling, error test, step accept -----
  P3D_FENCE();
  if (L.phase == kNewton && L.nres == 2) {
    if (!L.jac_current) {
      L.need_jac = 1;  // new Jacobian at the predictor, same step size
    } else {
      L.h *= 0.5;
      if (!(L.h >= L.hmin)) {  // also ends the span on a NaN step size
        L.status = -1;
        L.phase = kSpanEnd;
      } else {
        L.pend *= 0.5;
        L.n_equal = 0;
        L.lu_valid = 0;
        L.phase = kAttempt;
      }
    }
  } else if (L.phase == kNewton && L.nres == 1) {
    const double ecq = ndf_err(L.order);
    const double en = rms_sum(L.err_sq);
    if (en > 1.0) {
      const double factor = fmax(0.2, 0.9 * pow(en, -1.0 / (L.order + 1)));
      L.h *= factor;
      L.pend *= factor;
      L.n_equal = 0;
      L.lu_valid = 0;
      L.phase = kAttempt;
      if (!(L.h > 0.0)) {  // step size underflow or NaN: the span fails
        L.status = -1;
        L.phase = kSpanEnd;
      }
    } else {
      L.t = L.tn;
      L.st_steps++;
      L.jac_current = 0;
      const int q = L.order;
      // D[q+2] = d - D[q+1]; D[q+1] = d; D[j] += D[j+1] for j = q..0; and the
      // order-selection norms on the updated differences
      double s_e0 = 0.0, s_em = 0.0, s_ep = 0.0;
      const double ecm = q > 1 ? ndf_err(q - 1) : 0.0, ecp = q < kMaxOrder ? ndf_err(q + 1) : 0.0;
      for (int c0 = 0; c0 < kNS; c0 += kChunk4) {
        double Dv[kChunk4][kNumD], dvv[kChunk4];
#pragma unroll
        for (int u = 0; u < kChunk4; ++u) {
          dvv[u] = c0 + u < kNS ? (double)m.dd(c0 + u) : 0.0;
#pragma unroll
          for (int j = 0; j < kNumD; ++j)
            Dv[u][j] = (c0 + u < kNS && j <= q + 2) ? (double)m.D(j, c0 + u) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kChunk4; ++u) {
          if (c0 + u < kNS) {
            const double dv = dvv[u];
#pragma unroll
            for (int j = kNumD - 1; j >= 2; --j)
              if (j == q + 2) Dv[u][j] = dv - Dv[u][j - 1];
#pragma unroll
            for (int j = kNumD - 1; j >= 1; --j)
              if (j == q + 1) Dv[u][j] = dv;
#pragma unroll
            for (int j = kNumD - 2; j >= 0; --j)
              if (j <= q) Dv[u][j] += Dv[u][j + 1];
#pragma unroll
            for (int j = 0; j < kNumD; ++j)
              if (j <= q + 2) m.D(j, c0 + u) = Dv[u][j];
            const double sc = o.atol + o.rtol * fabs(Dv[u][0]);
            const double e = ecq * dv;
            s_e0 += (e / sc) * (e / sc);
            double dq = 0.0, dq2 = 0.0;
#pragma unroll
            for (int j = 0; j < kNumD; ++j) {
              if (j == q) dq = Dv[u][j];
              if (j == q + 2) dq2 = Dv[u][j];
            }
            const double em_ = ecm * dq / sc, ep_ = ecp * dq2 / sc;
            s_em += em_ * em_;
            s_ep += ep_ * ep_;
          }
        }
      }
      L.n_equal++;
      if (L.t >= L.t1) {
        L.phase = kSpanEnd;
      } else if (L.n_equal < q + 1) {
        L.phase = kStep;
      } else {
        const double e0 = rms_sum(s_e0);
        const double em = q > 1 ? rms_sum(s_em) : INFINITY;
        const double ep = q < kMaxOrder ? rms_sum(s_ep) : INFINITY;
        const double fm = q > 1 ? (em > 0.0 ? pow(em, -1.0 / (double)q) : INFINITY) : 0.0;
        const double fz = e0 > 0.0 ? pow(e0, -1.0 / (double)(q + 1)) : INFINITY;
        const double fp =
            q < kMaxOrder ? (ep > 0.0 ? pow(ep, -1.0 / (double)(q + 2)) : INFINITY) : 0.0;
        int dord = 0;
        double best = fz;
        if (fm > best) { best = fm; dord = -1; }
        if (fp > best) { best = fp; dord = 1; }
        L.order = q + dord;
        const double factor = fmin(10.0, 0.9 * best);
        L.h *= factor;
        L.pend *= factor;
        L.n_equal = 0;
        L.lu_valid = 0;
        L.phase = kStep;
      }
    }
  }
  // ---- end of a span: early-stop test, next checkpoint ---------------------
  P3D_FENCE();
  if (L.phase == kSpanEnd) {
    if (L.status == 0) {
      double yo[kNS], yn[kNS];
#pragma unroll
      for (int s = 0; s < kNS; ++s) {
        yo[s] = m.yc(s);
        yn[s] = L.trivial ? yo[s] : (double)m.D(0, s);
      }
      double rel = -1.0;
#pragma unroll
      for (int s = 0; s < kNS; ++s) {
        yn[s] = fmax(yn[s], 1e-30);
        if (o.early_stop &&
            (s == kRoles[0] || s == kRoles[1] || s == kRoles[3] || s == kRoles[4] ||
             s == kRoles[5] || s == kRoles[6])) {
          if (yo[s] > 1e-12) rel = fmax(rel, fabs(yn[s] - yo[s]) / fmax(yo[s], 1e-20));
        }
      }
#pragma unroll
      for (int s = 0; s < kNS; ++s) m.yc(s) = yn[s];
      if (o.early_stop && rel >= 0.0 && rel < o.stop_rtol && L.span > 3) L.conv = 1;
    }
    L.span++;
    L.phase = L.conv ? kFinish : kSpan;
  }
}

}  // namespace lane
}  // namespace p3d
