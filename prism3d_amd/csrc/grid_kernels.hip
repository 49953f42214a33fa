// Whole-grid kernels around the per-cell solve: multi-ray and point-source
// FUV columns, self-shielding + cosmic-ray attenuation, dust state and its
// slow evolution, the explicit-chemistry rate precompute, the thermal context
// and the heating/cooling diagnostics.
// All are HBM-streaming kernels over structure-of-arrays float64 buffers.
//
// Upstream (paths relative to the upstream repository):
//   columns      prism3d/radiative_transfer/fuv_rt_3d.py:215-392, fuv_rt.py:19-57
//   point source prism3d/radiative_transfer/fuv_rt_3d.py:74-212
//   shielding    prism3d/radiative_transfer/shielding.py:17-59, 136-171
//   cosmic rays  prism3d/cosmic_rays/attenuation.py:18-113 (model 'M')
//   dust state   prism3d/grains/themis_tables.py:278-354
//   dust evolve  prism3d/solver_3d.py:465-507
//   euler rates  prism3d/solver_3d.py:549-557
//   thermal ctx  prism3d/solver_3d.py:693-731
//   diagnostics  prism3d/solver_3d.py:898-1050
#include <algorithm>

#include "p3d_common.h"

namespace p3d {

// ===========================================================================
// Columns.  Along a ray whose dominant axis is `ax`, every cell's column is
// ds * (the layer values summed from the entry face up to and including the
// cell, minus half the cell's own layer), ds = cell / max(|d_ax|, 0.01).
// The sums from the low face (lo) and from the high face (hi) are computed
// once per axis and shared by every ray with that dominant axis; both run
// sequentially in the same order as a cumulative sum from that face.
// Workspace: F[axis][face][quantity][cell], 18 n doubles.
// ===========================================================================

// Prefix sums along axis 0 or 1 (stride > 1): one thread per line and face,
// consecutive threads on consecutive inner indices (coalesced).
__global__ void __launch_bounds__(256)
prefix_strided(int64_t nx, int64_t ny, int64_t nz, int axis, const double *__restrict__ dens,
               const double *__restrict__ xh2, const double *__restrict__ xco,
               double *__restrict__ F /* 2 faces x 3 quantities x n */) {
  const int64_t n = nx * ny * nz;
  const int64_t nlines = axis == 0 ? ny * nz : nx * nz;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * nlines) return;
  const int face = t >= nlines ? 1 : 0;
  const int64_t line = t - face * nlines;
  int64_t len, stride, base;
  if (axis == 0) {
    len = nx; stride = ny * nz; base = line;
  } else {
    const int64_t ix = line / nz, iz = line - ix * nz;
    len = ny; stride = nz; base = ix * ny * nz + iz;
  }
  double *Fq = F + (int64_t)face * 3 * n;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int64_t k = 0; k < len; ++k) {
    const int64_t i = base + (face ? (len - 1 - k) : k) * stride;
    const double d = dens[i];
    a += d;
    b += d * xh2[i] * 2.0;
    c += d * xco[i];
    Fq[i] = a;
    Fq[n + i] = b;
    Fq[2 * n + i] = c;
  }
}

// Prefix sums along the contiguous axis 2: one wavefront per line and face,
// 64 cells per chunk read coalesced, scanned with shuffles, carried across
// chunks (the within-chunk association differs from a strictly sequential
// sum by rounding only).
__global__ void __launch_bounds__(256)
prefix_contig(int64_t nx, int64_t ny, int64_t nz, const double *__restrict__ dens,
              const double *__restrict__ xh2, const double *__restrict__ xco,
              double *__restrict__ F) {
  const int64_t n = nx * ny * nz;
  const int lane = threadIdx.x & 63;
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nlines = nx * ny;
  if (w >= 2 * nlines) return;
  const int face = w >= nlines ? 1 : 0;
  const int64_t base = (w - face * nlines) * nz;
  double *Fq = F + (int64_t)face * 3 * n;
  double ca = 0.0, cb = 0.0, cc = 0.0;
  for (int64_t k0 = 0; k0 < nz; k0 += 64) {
    const int64_t k = k0 + lane;
    const int64_t i = base + (face ? (nz - 1 - k) : k);
    double a = 0.0, b = 0.0, c = 0.0;
    if (k < nz) {
      const double d = dens[i];
      a = d;
      b = d * xh2[i] * 2.0;
      c = d * xco[i];
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      double ta = __shfl_up(a, off, 64), tb = __shfl_up(b, off, 64), tc = __shfl_up(c, off, 64);
      if (lane >= off) { a += ta; b += tb; c += tc; }
    }
    a += ca; b += cb; c += cc;
    if (k < nz) {
      Fq[i] = a;
      Fq[n + i] = b;
      Fq[2 * n + i] = c;
    }
    ca = bcast(a, 63); cb = bcast(b, 63); cc = bcast(c, 63);
  }
}

constexpr int kMaxRays = 128;
struct RaySet {  // passed by value (kernel argument, < 4 KiB)
  int n;
  double g0w[kMaxRays];        // G0_ray * w / (4 pi)
  double ds[kMaxRays];         // path length per cell
  signed char ax[kMaxRays];    // dominant axis
  signed char from_low[kMaxRays];  // 1: ray enters at index 0 of its axis
};

__global__ void __launch_bounds__(256)
ray_combine(int64_t nx, int64_t ny, int64_t nz, const double *__restrict__ dens,
            const double *__restrict__ xh2, const double *__restrict__ xco,
            const double *__restrict__ F, const RaySet R, double av_per_nh,
            double *__restrict__ g0, double *__restrict__ av, double *__restrict__ nh,
            double *__restrict__ nh2, double *__restrict__ nco) {
  const int64_t n = nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = dens[i];
  const double l[3] = {d, d * xh2[i] * 2.0, d * xco[i]};
  // half-cell-corrected sums per (axis, face, quantity)
  double S[3][2][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int q = 0; q < 3; ++q) S[a][f][q] = F[((a * 2 + f) * 3 + q) * n + i] - 0.5 * l[q];
  double G = 0.0, avm = 1e10, bh = 0.0, bh2 = 0.0, bco = 0.0;
  const int nr = R.n;
  for (int r = 0; r < nr; ++r) {
    const int a = R.ax[r];
    const int f = R.from_low[r] ? 0 : 1;
    const double ds = R.ds[r];
    double s0, s1, s2;
    // select without dynamic register indexing
    if (a == 0) { s0 = f ? S[0][1][0] : S[0][0][0]; s1 = f ? S[0][1][1] : S[0][0][1]; s2 = f ? S[0][1][2] : S[0][0][2]; }
    else if (a == 1) { s0 = f ? S[1][1][0] : S[1][0][0]; s1 = f ? S[1][1][1] : S[1][0][1]; s2 = f ? S[1][1][2] : S[1][0][2]; }
    else { s0 = f ? S[2][1][0] : S[2][0][0]; s1 = f ? S[2][1][1] : S[2][0][1]; s2 = f ? S[2][1][2] : S[2][0][2]; }
    const double cnh = fmax(ds * s0, 0.0);
    const double cav = cnh * av_per_nh;
    G += R.g0w[r] * exp(-1.8 * cav);
    if (cav < avm) {
      avm = cav;
      bh = cnh;
      bh2 = fmax(ds * s1, 0.0);
      bco = fmax(ds * s2, 0.0);
    }
  }
  g0[i] = G; av[i] = avm; nh[i] = bh; nh2[i] = bh2; nco[i] = bco;
}

// Deterministic mean of n values in two passes over a fixed partition: block
// b of mean_partial sums the contiguous chunk [b*chunk, (b+1)*chunk) (strided
// per thread, then a fixed-shape tree) into part[b]; mean_final reduces the
// at most kMeanParts partials with a fixed tree and writes out[0].  The
// result depends only on n, never on scheduling.
constexpr int kMeanParts = 1024;

__global__ void __launch_bounds__(256) mean_partial(int64_t n, int64_t chunk,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ part) {
  __shared__ double s[256];
  const int64_t b0 = (int64_t)blockIdx.x * chunk;
  const int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
  double acc = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) acc += x[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(1024) mean_final(int nb, int64_t n,
                                                   const double *__restrict__ part,
                                                   double *__restrict__ out) {
  __shared__ double s[kMeanParts];
  s[threadIdx.x] = (int)threadIdx.x < nb ? part[threadIdx.x] : 0.0;
  __syncthreads();
  for (int w = kMeanParts / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0] / (double)n;
}

// Padovani+ (2018) piecewise fits in x = log10(N_H / 1e20) (attenuation.py:63-113)
__device__ __forceinline__ double pad_fit(double x, double a0, double b1, double b2, double c0,
                                          double c1, double d0) {
  double lz = x < -1.0 ? a0 : (x < 2.0 ? a0 + b1 * x + b2 * x * x : (x < 5.0 ? c0 + c1 * (x - 2.0) : d0));
  return exp10(lz);
}

__global__ void __launch_bounds__(256)
shield_cr_kernel(int64_t n, const double *__restrict__ nh2c, const double *__restrict__ ncoc,
                 const double *__restrict__ nhc, const double *__restrict__ t_mean,
                 double zeta0, double *__restrict__ fsh_h2, double *__restrict__ fsh_co,
                 double *__restrict__ zeta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // H2: Draine & Bertoldi (1996) eq. 37 with b from the grid-mean T
  const double b = sqrt(2.0 * kK * t_mean[0] / (2.0 * kMH)) / 1e5;
  const double x = nh2c[i] / 5.0e14;
  const double sq = sqrt(1.0 + x);
  const double q = 1.0 + x / b;
  fsh_h2[i] = fmin(fmax(0.965 / (q * q) + 0.035 / sq * exp(-8.5e-4 * sq), 0.0), 1.0);
  // CO: three-term fit of shielding.py:136-171
  const double nco = fmax(ncoc[i], 1.0), nh2 = fmax(nh2c[i], 1.0);
  const double lco = log(nco), lh2 = log(nh2);
  const double f = exp(-6.428e-8 * exp(0.53 * lco)) * exp(-3.0e-12 * exp(0.495 * lh2)) /
                   (1.0 + 1e-5 * exp(0.6 * lco) * exp(-8e-12 * exp(0.48 * lh2)));
  fsh_co[i] = fmin(fmax(f, 1e-10), 1.0);
  // cosmic rays, model M = sqrt(L * H) scaled to zeta0
  const double xc = log10(fmax(nhc[i], 1e18) / 1e20);
  const double zl = pad_fit(xc, -15.7, -0.4, -0.1, -16.3, -0.8, -18.7);
  const double zh = pad_fit(xc, -15.0, -0.35, -0.08, -15.6, -0.7, -17.7);
  zeta[i] = fmax(sqrt(zl * zh) * (zeta0 / 2e-16), 1e-22);
}

__global__ void __launch_bounds__(256)
dust_state_kernel(int64_t n, const double *__restrict__ g0, const double *__restrict__ av,
                  const double *__restrict__ nh, const double *__restrict__ tg,
                  const double *__restrict__ fnano, double s_nano, double s_lc, double s_sil,
                  double *__restrict__ t_dust, double *__restrict__ r_h2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double G = g0[i], A = av[i], n_h = nh[i], T = tg[i];
  // dust temperature (themis_tables.py:325-354)
  double F = fmax(1.6e-3 * G * exp(-1.8 * A) + 1.6e-3 * 0.5 * exp(-A), 1e-30);
  double Td = fmin(fmax(16.4 * exp(0.167 * log(F / 1.6e-3)), 5.0), 200.0);
  if (n_h > 1e5) {
    const double fc = fmin(fmax(n_h / 1e7, 0.0), 1.0);
    Td = Td * (1.0 - fc) + T * fc;
  }
  t_dust[i] = Td;
  // H2 formation rate coefficient (themis_tables.py:278-319)
  const double vH = sqrt(8.0 * kK * T / (kPi * kMH));
  const double S = 1.0 / (1.0 + 0.04 * sqrt(T + Td) + 0.002 * T + 8e-6 * T * T);
  const double eS = Td < 14.0 ? 1.0 : (Td < 25.0 ? exp(-(Td - 14.0) / 5.0) : 0.01);
  const double eC = fmax(Td < 50.0 ? 0.8 : (Td < 400.0 ? 0.8 * exp(-(Td - 50.0) / 200.0) : 0.01), 0.005);
  r_h2[i] = 0.5 * vH * S * (eC * s_nano * fnano[i] + eC * s_lc + eS * s_sil);
}

__global__ void __launch_bounds__(256)
euler_rates_kernel(int64_t n, const double *__restrict__ nh, const double *__restrict__ g0,
                   const double *__restrict__ av, const double *__restrict__ tg,
                   const double *__restrict__ zeta, const double *__restrict__ fh2,
                   const double *__restrict__ fco, const double *__restrict__ rh2,
                   double *__restrict__ kpd, double *__restrict__ kci, double *__restrict__ acp,
                   double *__restrict__ kco, double *__restrict__ rh2h, double *__restrict__ rcr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double G = g0[i], A = av[i], z = zeta[i];
  kpd[i] = 4.43e-11 * G * exp(-3.74 * A) * fh2[i];
  kci[i] = 2.56e-10 * G * exp(-3.02 * A);
  acp[i] = 4.67e-12 * pow(tg[i] / 300.0, -0.6);
  kco[i] = 1.71e-10 * G * exp(-3.53 * A) * fco[i] + 6.0 * z;
  rh2h[i] = rh2[i] * nh[i] * 0.5;
  rcr[i] = 2.5 * z * 0.1;
}

__global__ void __launch_bounds__(256)
thermal_ctx_kernel(int64_t n, const double *__restrict__ nh, const double *__restrict__ g0,
                   const double *__restrict__ av, const double *__restrict__ fh2,
                   const double *__restrict__ zeta, const double *__restrict__ td,
                   const double *__restrict__ gpe, const double *__restrict__ rh2,
                   const double *__restrict__ xhi, const double *__restrict__ xh2,
                   const double *__restrict__ xe, const double *__restrict__ xco,
                   const double *__restrict__ xcp, const double *__restrict__ xo,
                   const double *__restrict__ xc, double *__restrict__ gf,
                   double *__restrict__ gg, double *__restrict__ ne, double *__restrict__ nhi,
                   double *__restrict__ nh2, double *__restrict__ nco, double *__restrict__ ncp,
                   double *__restrict__ no, double *__restrict__ nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double n_h = nh[i], G = g0[i], A = av[i], XH2 = xh2[i], XHI = xhi[i];
  const double uv = G * exp(-3.74 * A) * fh2[i];
  double g = gpe[i];
  g += 3.3e-11 * uv * XH2 * n_h * (0.4 * kEV);
  g += rh2[i] * XHI * (n_h * n_h) * (1.5 * kEV);
  g += 3.0e-10 * uv * XH2 * n_h * (2.0 * kEV) * (n_h / (n_h + 3e4));
  g += zeta[i] * (XHI + 2.0 * XH2) * n_h * (20.0 * kEV);
  g += 3.0e-10 * G * exp(-3.0 * A) * xc[i] * n_h * kEV;
  gf[i] = g;
  const double mu = XH2 > 0.5 ? 2.0 : 1.0;
  gg[i] = 2.0 * 0.35 * 1e-21 * (n_h * n_h) * sqrt(8.0 * kK / (kPi * mu * kMH)) * kK;
  ne[i] = xe[i] * n_h;
  nhi[i] = XHI * n_h;
  nh2[i] = XH2 * n_h;
  nco[i] = xco[i] * n_h;
  ncp[i] = xcp[i] * n_h;
  no[i] = xo[i] * n_h;
  nc[i] = xc[i] * n_h;
  (void)td;
}

struct LineSet {
  double p[25];
};

__global__ void __launch_bounds__(256)
heat_cool_kernel(int64_t n, const double *__restrict__ tg, const double *__restrict__ nh,
                 const double *__restrict__ g0, const double *__restrict__ av,
                 const double *__restrict__ zeta, const double *__restrict__ fh2,
                 const double *__restrict__ xe, const double *__restrict__ xhi,
                 const double *__restrict__ xh2, const double *__restrict__ xc,
                 const double *__restrict__ xcp, const double *__restrict__ xo,
                 const double *__restrict__ xco, const double *__restrict__ td,
                 const double *__restrict__ gpe, const double *__restrict__ rh2, LineSet Ls,
                 double *__restrict__ gam, double *__restrict__ lam) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double T = tg[i], n_h = nh[i], G = g0[i], A = av[i], z = zeta[i];
  const double XE = xe[i], XHI = xhi[i], XH2 = xh2[i], XC = xc[i], XCP = xcp[i], XO = xo[i],
               XCO = xco[i], Td = td[i];
  // heating (solver_3d.py:898-934)
  double Gm = gpe[i];
  Gm += 3.3e-11 * G * exp(-3.74 * A) * fh2[i] * XH2 * n_h * 0.4 * kEV;
  Gm += rh2[i] * XHI * n_h * n_h * 1.5 * kEV;
  Gm += 3.0e-10 * G * exp(-3.74 * A) * fh2[i] * XH2 * n_h * 2.0 * kEV * (n_h / (n_h + 3e4));
  Gm += z * (XHI * n_h + 2.0 * XH2 * n_h) * 20.0 * kEV;
  const double mu = XH2 > 0.5 ? 2.0 : 1.0;
  const double vth = sqrt(8.0 * kK * T / (kPi * mu * kMH));
  Gm += 2.0 * 0.35 * 1e-21 * n_h * n_h * vth * kK * (Td - T);
  Gm += 3.0e-10 * G * exp(-3.0 * A) * XC * n_h * 1.0 * kEV;
  gam[i] = Gm;
  // cooling (solver_3d.py:936-1050)
  const double n_e = XE * n_h, n_hi = XHI * n_h, n_h2 = XH2 * n_h;
  const double sT = sqrt(T), t01 = T / 100.0;
  double L = 0.0;
  {
    const double *p = Ls.p;
    const double bz = exp(-p[2] / T);
    const double q = 8.63e-6 / (p[4] * sT) * 2.15 * bz * n_e + 8.0e-10 * pow(t01, 0.07) * n_hi +
                     4.9e-10 * pow(t01, 0.12) * n_h2;
    const double rul = p[1] + q, rlu = (p[3] / p[4]) * q * bz;
    L += XCP * n_h * rlu / (rlu + rul) * p[1] * kH * p[0];
  }
  {
    const double *p = Ls.p + 5;
    const double q = 9.2e-12 * pow(T, 0.67) * n_hi + 4.5e-12 * pow(T, 0.64) * n_h2 +
                     1.4e-8 * pow(t01, 0.39) * n_e;
    const double rlu = (p[3] / p[4]) * q * exp(-p[2] / T), rul = p[1] + q;
    L += XO * n_h * rlu / (rlu + rul) * p[1] * kH * p[0];
  }
  {
    const double *p = Ls.p + 10;
    const double q = 4.0e-12 * pow(T, 0.60) * n_hi + 2.0e-12 * pow(T, 0.58) * n_h2 +
                     5.0e-9 * pow(t01, 0.39) * n_e;
    const double rlu = (p[3] / p[4]) * q * exp(-p[2] / T), rul = p[1] + q;
    L += XO * n_h * rlu / (rlu + rul) * p[1] * kH * p[0];
  }
  for (int k = 0; k < 2; ++k) {
    const double *p = Ls.p + 15 + 5 * k;
    const double q = 1.0e-10 * pow(t01, 0.3) * n_hi + 5.0e-11 * pow(t01, 0.3) * n_h2 +
                     2.0e-7 * pow(t01, -0.5) * n_e;
    const double rlu = (p[3] / p[4]) * q * exp(-p[2] / T), rul = p[1] + q;
    L += XC * n_h * rlu / (rlu + rul) * p[1] * kH * p[0];
  }
  {
    const double n_co = XCO * n_h, B = 57.635968e9, mud = 0.1098e-18;
    for (int j = 1; j <= 40; ++j) {
      const double J = j, freq = 2.0 * B * J, Tu = kH * B * J * (J + 1.0) / kK;
      if (!(Tu < 5.0 * T)) break;
      const double Aco = (64.0 * kPi * kPi * kPi * kPi * freq * freq * freq * mud * mud * J) /
                         (3.0 * kH * kC * kC * kC * (2.0 * J + 1.0));
      const double gu = 2.0 * J + 1.0, gl = 2.0 * (J - 1.0) + 1.0;
      const double g = 3.3e-11 * sqrt(t01) * (1.0 + 0.1 * J);
      const double rlu = (gu / gl) * g * n_h2 * exp(-Tu / T), rul = Aco + g * n_h2;
      L += n_co * rlu / (rlu + rul) * Aco * kH * freq;
    }
  }
  if (T >= 100.0) {
    double x = fmin(fmax(log10(fmax(T, 100.0)), 2.0), 4.5);
    const double x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x, x6 = x5 * x;
    const double lh = -24.311 + 3.5692 * x - 11.332 * x2 + 15.738 * x3 - 10.581 * x4 + 3.5803 * x5 - 0.48365 * x6;
    const double lh2 = -24.311 + 4.6585 * x - 14.272 * x2 + 19.779 * x3 - 13.255 * x4 + 4.4840 * x5 - 0.60504 * x6;
    const double l0 = exp10(lh) * n_hi + exp10(lh2) * n_h2, llte = exp10(-19.703 + 0.5 * x);
    L += n_h2 * l0 * llte / fmax(l0 + llte, 1e-50);
  }
  if (T >= 3000.0) L += n_hi * n_e * (2.41e-6 / sT * 0.486 * exp(-1.18e5 / T)) * 10.2 * kEV;
  L += fmax(2.0 * 0.35 * 1e-21 * n_h * n_h * vth * kK * (T - Td), 0.0);
  L += XCP * n_h * n_e * 4.67e-12 * pow(T / 300.0, -0.6) * kK * T;
  lam[i] = L;
}

// ===========================================================================
// Point-source columns (fuv_rt_3d.py:74-212).  The star's cell is
// (position / cell) clipped to the grid and truncated, per axis.  Along an
// axis, a cell's column is the sum of the layers between the star's plane
// (included) and the cell (excluded), accumulated outward from the star's
// plane, which is the upstream cumulative-sum order.  A cell uses only the
// column along its dominant axis (largest |displacement| from the star; ties
// go to x, then y), so one thread per (axis, line, direction) walks outward
// from the star's plane carrying the running sums and writes the final G0,
// A_V and columns of exactly the cells whose dominant axis it walks.  Every
// cell is written once; every cell is read by three walks (one per axis).
// Arithmetic is uncontracted so that sums and products round as upstream.
// ===========================================================================
struct PointSrc {
  double sx, sy, sz;         // star position [cm]
  double cell, l_fuv;        // cell edge [cm], FUV luminosity [erg/s]
  double av_per_nh, g0_to_flux;
  int64_t s[3];              // star cell per axis
};

// dominant axis of a displacement, upstream tie rule
__device__ __forceinline__ int ps_dominant(double ax, double ay, double az) {
  if (ax >= ay && ax >= az) return 0;
  if (ay > ax && ay >= az) return 1;
  return 2;
}

__device__ __forceinline__ void ps_finish(const PointSrc &P, double dx, double dy, double dz,
                                          double ddom, double c0, double c1, double c2,
                                          double l0, double l1, double l2, int64_t i,
                                          double *__restrict__ g0, double *__restrict__ av,
                                          double *__restrict__ nh, double *__restrict__ nh2,
                                          double *__restrict__ nco) {
#pragma clang fp contract(off)
  const double half = 0.5 * P.cell;
  const double r = fmax(sqrt(dx * dx + dy * dy + dz * dz), half);
  const double pc = fmin(r / fmax(ddom, half), 3.0);
  const double h = fmax(c0 * pc - 0.5 * l0 * pc, 0.0);
  nh[i] = h;
  nh2[i] = fmax(c1 * pc - 0.5 * l1 * pc, 0.0);
  nco[i] = fmax(c2 * pc - 0.5 * l2 * pc, 0.0);
  const double a = h * P.av_per_nh;
  av[i] = a;
  g0[i] = P.l_fuv / (4.0 * kPi * (r * r) * P.g0_to_flux) * exp(-1.8 * a);
}

__global__ void __launch_bounds__(256)
point_source_kernel(int64_t nx, int64_t ny, int64_t nz, const double *__restrict__ dens,
                    const double *__restrict__ xh2, const double *__restrict__ xco,
                    const PointSrc P, double *__restrict__ g0, double *__restrict__ av,
                    double *__restrict__ nh, double *__restrict__ nh2,
                    double *__restrict__ nco) {
#pragma clang fp contract(off)
  // thread ranges: axis 0 lines (y, z), axis 1 lines (x, z), axis 2 lines (x, y), each twice
  // (outward toward high indices, then toward low indices)
  const int64_t n0 = ny * nz, n1 = nx * nz, n2 = nx * ny;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int axis;
  int64_t nl;
  if (t < 2 * n0) { axis = 0; nl = n0; }
  else if ((t -= 2 * n0) < 2 * n1) { axis = 1; nl = n1; }
  else if ((t -= 2 * n1) < 2 * n2) { axis = 2; nl = n2; }
  else return;
  const int up = t < nl ? 1 : 0;
  const int64_t line = up ? t : t - nl;
  int64_t len, stride, base, ix = 0, iy = 0, iz = 0;
  if (axis == 0) {
    len = nx; stride = ny * nz; base = line; iy = line / nz; iz = line - iy * nz;
  } else if (axis == 1) {
    ix = line / nz; iz = line - ix * nz;
    len = ny; stride = nz; base = ix * ny * nz + iz;
  } else {
    ix = line / ny; iy = line - ix * ny;
    len = nz; stride = 1; base = line * nz;
  }
  const int64_t s = P.s[axis];
  // displacement components of the two fixed coordinates of this line
  const double cell = P.cell;
  const double dxf = ((double)ix + 0.5) * cell - P.sx;
  const double dyf = ((double)iy + 0.5) * cell - P.sy;
  const double dzf = ((double)iz + 0.5) * cell - P.sz;
  const double sa = axis == 0 ? P.sx : (axis == 1 ? P.sy : P.sz);
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;
  auto layer = [&](int64_t i, double &l0, double &l1, double &l2) {
#pragma clang fp contract(off)
    const double d = dens[i];
    l0 = d * cell;
    l1 = xh2 ? d * xh2[i] * 2.0 * cell : 0.0 * cell;
    l2 = xco ? d * xco[i] * cell : 0.0 * cell;
  };
  auto visit = [&](int64_t k, double l0, double l1, double l2) {
#pragma clang fp contract(off)
    const double dk = ((double)k + 0.5) * cell - sa;
    const double dx = axis == 0 ? dk : dxf, dy = axis == 1 ? dk : dyf, dz = axis == 2 ? dk : dzf;
    const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
    if (ps_dominant(ax, ay, az) != axis) return;
    ps_finish(P, dx, dy, dz, axis == 0 ? ax : (axis == 1 ? ay : az), c0, c1, c2, l0, l1, l2,
              base + k * stride, g0, av, nh, nh2, nco);
  };
  if (up) {
    for (int64_t k = s; k < len; ++k) {
      double l0, l1, l2;
      layer(base + k * stride, l0, l1, l2);
      visit(k, l0, l1, l2);
      c0 += l0; c1 += l1; c2 += l2;
    }
  } else if (s > 0) {
    double p0, p1, p2;
    layer(base + s * stride, p0, p1, p2);
    for (int64_t k = s - 1; k >= 0; --k) {
      c0 += p0; c1 += p1; c2 += p2;
      double l0, l1, l2;
      layer(base + k * stride, l0, l1, l2);
      visit(k, l0, l1, l2);
      p0 = l0; p1 = l1; p2 = l2;
    }
  }
}

// ===========================================================================
// Slow grain evolution over one outer step of 10 kyr (solver_3d.py:465-507):
// nano-grain photo-destruction, replenishment and accretion, and band-gap
// relaxation toward its irradiation/density equilibrium.  In place on
// f_nano and E_g.
// ===========================================================================
__global__ void __launch_bounds__(256)
dust_evolve_kernel(int64_t n, const double *__restrict__ g0, const double *__restrict__ nh,
                   double *__restrict__ f_nano, double *__restrict__ e_g) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double yr = 3.15e7;
  const double dt = 1e4 * yr;
  const double G = g0[i], n_h = nh[i], fn = f_nano[i], eg = e_g[i];
  const double tau_dest = 1e4 * yr / fmax(G / 1e4, 0.01);
  const double f_dest = exp(-dt / tau_dest);
  const double f_rep = fmin(dt / (1e6 * yr), 0.1);
  const double f_acc = (G < 1.0 && n_h > 1e4) ? fmin(dt / (1e5 * yr * 1e4 / n_h), 0.1) : 0.0;
  f_nano[i] = fmin(fmax(fn * f_dest + (1.0 - fn) * f_rep + f_acc, 0.01), 2.0);
  const double tau_ar = 1e3 * yr / fmax(G, 0.01);
  const double eg_eq = G > 10.0 ? 0.01 : fmin(0.1 + 0.5 * log10(fmax(n_h, 1.0)), 2.5);
  e_g[i] = fmin(fmax(eg + (eg_eq - eg) * fmin(dt / tau_ar, 1.0), 0.01), 2.5);
}

}  // namespace p3d

using namespace p3d;

extern "C" {

int p3d_fuv_columns(int64_t nx, int64_t ny, int64_t nz, const double *density,
                    const double *x_h2, const double *x_co, double cell_size, int32_t n_rays,
                    const double *dirs3, const double *g0_rays, double av_per_nh, double *g0,
                    double *a_v, double *n_h, double *n_h2, double *n_co, double *workspace,
                    void *stream) {
  P3D_REQUIRE(nx > 0 && ny > 0 && nz > 0, "grid extents must be > 0");
  P3D_REQUIRE(n_rays >= 1 && n_rays <= kMaxRays, "n_rays must be in [1, %d] (got %d)", kMaxRays,
              n_rays);
  P3D_REQUIRE(dirs3 && g0_rays, "ray directions and ray G0 (host) are required");
  P3D_REQUIRE(density && x_h2 && x_co && g0 && a_v && n_h && n_h2 && n_co && workspace,
              "null device array");
  P3D_REQUIRE(cell_size > 0.0, "cell_size must be > 0");
  const int64_t n = nx * ny * nz;
  RaySet R;
  R.n = n_rays;
  for (int r = 0; r < n_rays; ++r) {
    const double d[3] = {dirs3[3 * r], dirs3[3 * r + 1], dirs3[3 * r + 2]};
    const double ad[3] = {fabs(d[0]), fabs(d[1]), fabs(d[2])};
    int dom = 0;  // first maximum, as numpy argmax
    if (ad[1] > ad[dom]) dom = 1;
    if (ad[2] > ad[dom]) dom = 2;
    R.ax[r] = (signed char)dom;
    R.from_low[r] = d[dom] < 0.0 ? 1 : 0;
    R.ds[r] = cell_size * (1.0 / fmax(ad[dom], 0.01));
    R.g0w[r] = g0_rays[r] * ((4.0 * kPi / n_rays) / (4.0 * kPi));
  }
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(prefix_strided, dim3(grid_for(2 * ny * nz, 256)), dim3(256), 0, s, nx, ny,
                     nz, 0, density, x_h2, x_co, workspace);
  hipLaunchKernelGGL(prefix_strided, dim3(grid_for(2 * nx * nz, 256)), dim3(256), 0, s, nx, ny,
                     nz, 1, density, x_h2, x_co, workspace + 6 * n);
  hipLaunchKernelGGL(prefix_contig, dim3(grid_for(2 * nx * ny * 64, 256)), dim3(256), 0, s, nx,
                     ny, nz, density, x_h2, x_co, workspace + 12 * n);
  hipLaunchKernelGGL(ray_combine, dim3(grid_for(n, 256)), dim3(256), 0, s, nx, ny, nz, density,
                     x_h2, x_co, workspace, R, av_per_nh, g0, a_v, n_h, n_h2, n_co);
  return launch_status("fuv_columns");
}

int p3d_mean(int64_t n, const double *x, double *out, double *workspace, void *stream) {
  P3D_REQUIRE(n > 0 && x && out && workspace,
              "p3d_mean: n > 0 and device pointers (x, out, workspace) required");
  const int64_t nb = std::min<int64_t>(kMeanParts, std::max<int64_t>(1, (n + 8191) / 8192));
  const int64_t chunk = (n + nb - 1) / nb;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(mean_partial, dim3((unsigned)nb), dim3(256), 0, s, n, chunk, x, workspace);
  hipLaunchKernelGGL(mean_final, dim3(1), dim3(kMeanParts), 0, s, (int)nb, n, workspace, out);
  return launch_status("mean_kernel");
}

int p3d_shield_cr(int64_t n, const double *n_h2_col, const double *n_co_col,
                  const double *n_h_col, const double *t_mean, double zeta0, double *fsh_h2,
                  double *fsh_co, double *zeta, void *stream) {
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(n_h2_col && n_co_col && n_h_col && t_mean && fsh_h2 && fsh_co && zeta,
              "null device array");
  hipLaunchKernelGGL(shield_cr_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     n, n_h2_col, n_co_col, n_h_col, t_mean, zeta0, fsh_h2, fsh_co, zeta);
  return launch_status("shield_cr_kernel");
}

int p3d_dust_state(int64_t n, const double *g0, const double *a_v, const double *n_h,
                   const double *t_gas, const double *f_nano, double sigma_nano,
                   double sigma_large_c, double sigma_sil, double *t_dust, double *r_h2,
                   void *stream) {
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(g0 && a_v && n_h && t_gas && f_nano && t_dust && r_h2, "null device array");
  hipLaunchKernelGGL(dust_state_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     n, g0, a_v, n_h, t_gas, f_nano, sigma_nano, sigma_large_c, sigma_sil, t_dust,
                     r_h2);
  return launch_status("dust_state_kernel");
}

int p3d_euler_rates(int64_t n, const double *n_h, const double *g0, const double *a_v,
                    const double *t_gas, const double *zeta, const double *fsh_h2,
                    const double *fsh_co, const double *r_h2, double *k_pd_base, double *k_ci,
                    double *alpha_cp, double *k_co_total, double *r_h2_half_nh,
                    double *r_cr_base, void *stream) {
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(n_h && g0 && a_v && t_gas && zeta && fsh_h2 && fsh_co && r_h2 && k_pd_base && k_ci &&
                  alpha_cp && k_co_total && r_h2_half_nh && r_cr_base,
              "null device array");
  hipLaunchKernelGGL(euler_rates_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     n, n_h, g0, a_v, t_gas, zeta, fsh_h2, fsh_co, r_h2, k_pd_base, k_ci, alpha_cp,
                     k_co_total, r_h2_half_nh, r_cr_base);
  return launch_status("euler_rates_kernel");
}

int p3d_thermal_ctx(int64_t n, const double *n_h, const double *g0, const double *a_v,
                    const double *fsh_h2, const double *zeta, const double *t_dust,
                    const double *gamma_pe, const double *r_h2, const double *x_hi,
                    const double *x_h2, const double *x_e, const double *x_co,
                    const double *x_cp, const double *x_o, const double *x_c,
                    double *gamma_fixed, double *gg_coeff, double *n_e, double *n_hi,
                    double *n_h2, double *n_co, double *n_cp_nh, double *n_o_nh,
                    double *n_c_nh, void *stream) {
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(n_h && g0 && a_v && fsh_h2 && zeta && gamma_pe && r_h2 && x_hi && x_h2 && x_e &&
                  x_co && x_cp && x_o && x_c && gamma_fixed && gg_coeff && n_e && n_hi && n_h2 &&
                  n_co && n_cp_nh && n_o_nh && n_c_nh,
              "null device array");
  hipLaunchKernelGGL(thermal_ctx_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     n, n_h, g0, a_v, fsh_h2, zeta, t_dust, gamma_pe, r_h2, x_hi, x_h2, x_e, x_co,
                     x_cp, x_o, x_c, gamma_fixed, gg_coeff, n_e, n_hi, n_h2, n_co, n_cp_nh, n_o_nh,
                     n_c_nh);
  return launch_status("thermal_ctx_kernel");
}

int p3d_heating_cooling(int64_t n, const double *t_gas, const double *n_h, const double *g0,
                        const double *a_v, const double *zeta, const double *fsh_h2,
                        const double *x_e, const double *x_hi, const double *x_h2,
                        const double *x_c, const double *x_cp, const double *x_o,
                        const double *x_co, const double *t_dust, const double *gamma_pe,
                        const double *r_h2, const double *cii5, const double *oi63_5,
                        const double *oi145_5, const double *ci10, double *gamma_out,
                        double *lambda_out, void *stream) {
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(cii5 && oi63_5 && oi145_5 && ci10, "line parameter vectors are required");
  P3D_REQUIRE(t_gas && n_h && g0 && a_v && zeta && fsh_h2 && x_e && x_hi && x_h2 && x_c && x_cp &&
                  x_o && x_co && t_dust && gamma_pe && r_h2 && gamma_out && lambda_out,
              "null device array");
  LineSet Ls;
  for (int k = 0; k < 5; ++k) {
    Ls.p[k] = cii5[k];
    Ls.p[5 + k] = oi63_5[k];
    Ls.p[10 + k] = oi145_5[k];
    Ls.p[15 + k] = ci10[k];
    Ls.p[20 + k] = ci10[5 + k];
  }
  hipLaunchKernelGGL(heat_cool_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     n, t_gas, n_h, g0, a_v, zeta, fsh_h2, x_e, x_hi, x_h2, x_c, x_cp, x_o, x_co,
                     t_dust, gamma_pe, r_h2, Ls, gamma_out, lambda_out);
  return launch_status("heat_cool_kernel");
}

int p3d_fuv_point_source(int64_t nx, int64_t ny, int64_t nz, const double *density,
                         const double *x_h2, const double *x_co, const double *star_pos,
                         double l_fuv, double cell_size, double av_per_nh, double g0_to_flux,
                         double *g0, double *a_v, double *n_h, double *n_h2, double *n_co,
                         void *stream) {
  P3D_REQUIRE(nx > 0 && ny > 0 && nz > 0, "grid extents must be > 0");
  P3D_REQUIRE(star_pos, "star position (host, 3 doubles) is required");
  P3D_REQUIRE(density && g0 && a_v && n_h && n_h2 && n_co, "null device array");
  P3D_REQUIRE(cell_size > 0.0, "cell_size must be > 0");
  P3D_REQUIRE(std::isfinite(star_pos[0]) && std::isfinite(star_pos[1]) &&
                  std::isfinite(star_pos[2]),
              "star position must be finite");
  PointSrc P;
  P.sx = star_pos[0]; P.sy = star_pos[1]; P.sz = star_pos[2];
  P.cell = cell_size; P.l_fuv = l_fuv; P.av_per_nh = av_per_nh; P.g0_to_flux = g0_to_flux;
  const int64_t ext[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {  // int(np.clip(pos / cell, 0, n - 1))
    const double q = fmin(fmax(star_pos[a] / cell_size, 0.0), (double)(ext[a] - 1));
    P.s[a] = (int64_t)q;
  }
  const int64_t threads = 2 * (ny * nz + nx * nz + nx * ny);
  hipLaunchKernelGGL(point_source_kernel, dim3(grid_for(threads, 256)), dim3(256), 0,
                     as_stream(stream), nx, ny, nz, density, x_h2, x_co, P, g0, a_v, n_h, n_h2,
                     n_co);
  return launch_status("point_source_kernel");
}

int p3d_dust_evolve(int64_t n, const double *g0, const double *n_h, double *f_nano, double *e_g,
                    void *stream) {
  P3D_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return P3D_OK;
  P3D_REQUIRE(g0 && n_h && f_nano && e_g, "null device array");
  hipLaunchKernelGGL(dust_evolve_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     n, g0, n_h, f_nano, e_g);
  return launch_status("dust_evolve_kernel");
}

}  // extern "C"
