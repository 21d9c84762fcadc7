// tcsem_host.cpp -- host runtime of libtcsem.so: the C-ABI of include/tcsem.h.
//
// Owns the reference's file surface (FDTD_setup.in, Cond.bin, Receiver.in, Transmitter.in,
// Tx<id>.bout), the host setup it restates (grid, regular/irregular index maps, Gaussian source
// schedule, receiver/transmitter cell search), the precomputed operators of the air boundary
// (spline window weights, FFT plans, wavenumber gains), device memory, and the step driver of
// main.cu:195-298 with the EFInterp2Rx sampling state machine.  All numerics of a step run in the
// gfx950 kernels of tcsem_kernels.hip and tcsem_upcont.hip; there is no CPU fallback.
#include "../../include/tcsem.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "tcsem_internal.h"

using namespace tcsem;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(TCSEM_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

// ---- host setup restatements --------------------------------------------------------------------
struct HostGrid {
  tcsem_grid_info info{};
  int AL = 0, AR = 0, AY = 0;
  double mindx = 0;
  std::vector<double> dx, dy, dz, X, Y, Z, Xn, Yn;
  std::vector<int> xN2o, yN2o, xo2N, yo2N, xBx2N, yBy2N;
};

void coords(const tcsem_setup &s, HostGrid &g) {  // ModelingSetup.cpp:154-187
  const int L = g.info.L, M = g.info.M, N = g.info.N;
  int nc = (s.Nx % 2 == 0) ? s.Nx / 2 + g.AL : (s.Nx - 1) / 2 + g.AL;
  g.X[nc] = (s.Nx % 2 == 0) ? 0.5 * g.dx[nc] + s.x0 : s.x0;
  for (int i = nc + 1; i < L; ++i) g.X[i] = g.X[i - 1] + 0.5 * g.dx[i - 1] + 0.5 * g.dx[i];
  for (int i = nc - 1; i >= 0; --i) g.X[i] = g.X[i + 1] - 0.5 * g.dx[i + 1] - 0.5 * g.dx[i];
  nc = (s.Ny % 2 == 0) ? s.Ny / 2 + g.AY : (s.Ny - 1) / 2 + g.AY;
  g.Y[nc] = (s.Ny % 2 == 0) ? 0.5 * g.dy[nc] + s.y0 : s.y0;
  for (int j = nc + 1; j < M; ++j) g.Y[j] = g.Y[j - 1] + 0.5 * g.dy[j - 1] + 0.5 * g.dy[j];
  for (int j = nc - 1; j >= 0; --j) g.Y[j] = g.Y[j + 1] - 0.5 * g.dy[j + 1] - 0.5 * g.dy[j];
  g.Z[0] = 0.5 * g.dz[0] + s.z0;
  for (int k = 1; k < N; ++k) g.Z[k] = g.Z[k - 1] + 0.5 * g.dz[k - 1] + 0.5 * g.dz[k];
  // regular air-plane widths (ModelingSetup.cpp:190-192)
  g.info.AnLx_New = (int)std::ceil((g.X[g.AL - 1] - g.X[0] + g.dx[g.AL - 1] / 2 + g.dx[0] / 2) / s.dxmin);
  g.info.AnRx_New = (int)std::ceil((g.X[L - 1] - g.X[L - g.AR] + g.dx[L - 1] / 2 + g.dx[L - g.AR] / 2) / s.dxmin);
  g.info.Any_New = (int)std::ceil((g.Y[g.AY - 1] - g.Y[0] + g.dy[g.AY - 1] / 2 + g.dy[0] / 2) / s.dymin);
  g.info.L1 = g.info.AnLx_New + g.info.AnRx_New + s.Nx;  // main.cu:102-103
  g.info.M1 = 2 * g.info.Any_New + s.Ny;
}

int build_grid(const tcsem_setup &s, HostGrid &g) {  // ModelingSetup.cpp:79-152
  if (!(s.dxmin > 0 && s.dymin > 0 && s.dzmin > 0 && s.arg1 > 0 && s.arg2 > 0))
    return fail(TCSEM_EINVAL, "cell sizes and growth factors must be positive");
  if (s.ALx1 < 0 || s.ALx2 < 0 || s.ARx1 < 0 || s.ARx2 < 0 || s.Ay1 < 0 || s.Ay2 < 0 || s.Az1 < 0 ||
      s.Az2 < 0 || s.Nz < 1)
    return fail(TCSEM_EINVAL, "negative zone count");
  g.AL = s.ALx1 + s.ALx2;
  g.AR = s.ARx1 + s.ARx2;
  g.AY = s.Ay1 + s.Ay2;
  // Zero padding makes the reference index X_Bzold[-1] (ModelingSetup.cpp:190-192).
  if (g.AL < 1 || g.AR < 1 || g.AY < 1)
    return fail(TCSEM_EINVAL, "each side needs at least one padding cell");
  if (s.Nx < kNPad || s.Ny < kNPad) return fail(TCSEM_EINVAL, "core needs at least %d cells per axis", kNPad);
  const int L = g.AL + g.AR + s.Nx, M = 2 * g.AY + s.Ny, N = s.Nz + s.Az1 + s.Az2;
  if (N < 3) return fail(TCSEM_EINVAL, "need at least 3 levels");
  if (g.AL + kNPad > L || g.AR + kNPad > L || g.AY + kNPad > M) return fail(TCSEM_EINVAL, "spline window exceeds grid");
  g.info.L = L;
  g.info.M = M;
  g.info.N = N;
  g.dx.assign(L, 0);
  g.dy.assign(M, 0);
  g.dz.assign(N, 0);
  g.X.assign(L, 0);
  g.Y.assign(M, 0);
  g.Z.assign(N, 0);
  for (int i = g.AL; i < s.Nx + g.AL; ++i) g.dx[i] = s.dxmin;
  for (int i = g.AL - 1; i >= s.ALx2; --i) g.dx[i] = s.dxmin * std::pow(s.arg1, double(g.AL - i));
  for (int i = s.ALx2 - 1; i >= 0; --i)
    g.dx[i] = s.dxmin * std::pow(s.arg1, double(s.ALx1)) * std::pow(s.arg2, double(s.ALx2 - i));
  for (int i = s.Nx + g.AL; i < s.Nx + g.AL + s.ARx1; ++i)
    g.dx[i] = s.dxmin * std::pow(s.arg1, double(i - g.AL - s.Nx + 1));
  for (int i = s.Nx + g.AL + s.ARx1; i < L; ++i)
    g.dx[i] = s.dxmin * std::pow(s.arg1, double(s.ARx1)) * std::pow(s.arg2, double(i - s.Nx - g.AL - s.ARx1 + 1));
  for (int j = g.AY; j < s.Ny + g.AY; ++j) g.dy[j] = s.dymin;
  for (int j = g.AY - 1; j >= s.Ay2; --j) g.dy[j] = g.dy[j + 1] * s.arg1;
  for (int j = s.Ay2 - 1; j >= 0; --j) g.dy[j] = g.dy[j + 1] * s.arg2;
  for (int j = s.Ny + g.AY; j < s.Ny + s.Ay2 + 2 * s.Ay1; ++j) g.dy[j] = g.dy[j - 1] * s.arg1;
  for (int j = s.Ny + s.Ay2 + 2 * s.Ay1; j < M; ++j) g.dy[j] = g.dy[j - 1] * s.arg2;
  for (int k = 0; k < s.Nz; ++k) g.dz[k] = s.dzmin;
  for (int k = s.Nz; k < s.Nz + s.Az1; ++k) g.dz[k] = g.dz[k - 1] * s.arg1;
  for (int k = s.Nz + s.Az1; k < N; ++k) g.dz[k] = g.dz[k - 1] * s.arg2;
  coords(s, g);
  g.mindx = std::min(s.dzmin, std::min(s.dxmin, s.dymin));  // main.cu:107
  return TCSEM_OK;
}

// Walks of ModelGridMapping.cpp:77-366: from a seed index, step towards the target until the
// bracket condition holds.
void build_maps(const tcsem_setup &s, HostGrid &g) {
  const int L0 = g.info.L, M0 = g.info.M, L1 = g.info.L1, M1 = g.info.M1;
  const int aLx = g.info.AnLx_New, aY = g.info.Any_New, Nx = s.Nx, Ny = s.Ny;
  g.Xn.assign(L1, 0);
  g.Yn.assign(M1, 0);
  for (int i = aLx; i < aLx + Nx; ++i) g.Xn[i] = g.X[i - aLx + g.AL];
  for (int i = aLx - 1; i >= 0; --i) g.Xn[i] = g.Xn[i + 1] - s.dxmin;
  for (int i = aLx + Nx; i < L1; ++i) g.Xn[i] = g.Xn[i - 1] + s.dxmin;
  for (int j = aY; j < aY + Ny; ++j) g.Yn[j] = g.Y[j - aY + g.AY];
  for (int j = aY - 1; j >= 0; --j) g.Yn[j] = g.Yn[j + 1] - s.dymin;
  for (int j = aY + Ny; j < M1; ++j) g.Yn[j] = g.Yn[j - 1] + s.dymin;

  // new -> old: index of the first old node >= the new point (core points coincide)
  auto new2old = [](const std::vector<double> &O, const std::vector<double> &Nw, int core0, int ncore,
                    int ocore0, std::vector<int> &out) {
    const int n0 = (int)O.size(), n1 = (int)Nw.size();
    out.assign(n1, 0);
    for (int i = core0; i < core0 + ncore; ++i) out[i] = i - core0 + ocore0;
    for (int i = core0 - 1; i >= 0; --i) {
      int j = out[i + 1];
      while (j > 0 && O[j - 1] >= Nw[i]) --j;
      out[i] = j;
    }
    for (int i = core0 + ncore; i < n1; ++i) {
      int j = out[i - 1] - 1;
      while (j < n0 - 1 && O[j + 1] < Nw[i]) ++j;
      out[i] = j + 1;
    }
  };
  new2old(g.X, g.Xn, aLx, Nx, g.AL, g.xN2o);
  new2old(g.Y, g.Yn, aY, Ny, g.AY, g.yN2o);
  // old (possibly shifted by -d/2, plus one end point) -> new
  auto old2new = [](const std::vector<double> &P, const std::vector<double> &Nw, int ocore0, int ncore,
                    int core0, std::vector<int> &out) {
    const int np = (int)P.size(), n1 = (int)Nw.size();
    out.assign(np, 0);
    for (int i = ocore0; i < ocore0 + ncore; ++i) out[i] = i - ocore0 + core0;
    for (int i = ocore0 - 1; i >= 0; --i) {
      int j = out[i + 1];
      while (j > 0 && Nw[j - 1] >= P[i]) --j;
      out[i] = j;
    }
    for (int i = ocore0 + ncore; i < np; ++i) {
      int j = out[i - 1];
      while (j < n1 - 1 && Nw[j + 1] < P[i]) ++j;
      out[i] = j + 1;
    }
  };
  old2new(g.X, g.Xn, g.AL, Nx, aLx, g.xo2N);
  old2new(g.Y, g.Yn, g.AY, Ny, aY, g.yo2N);
  std::vector<double> px(L0 + 1), py(M0 + 1);
  for (int i = 0; i < L0; ++i) px[i] = g.X[i] - 0.5 * g.dx[i];
  px[L0] = g.X[L0 - 1] + 0.5 * g.dx[L0 - 1];
  for (int j = 0; j < M0; ++j) py[j] = g.Y[j] - 0.5 * g.dy[j];
  py[M0] = g.Y[M0 - 1] + 0.5 * g.dy[M0 - 1];
  old2new(px, g.Xn, g.AL, Nx, aLx, g.xBx2N);
  old2new(py, g.Yn, g.AY, Ny, aY, g.yBy2N);
}

// SourceWave.cpp:8-58; dt[2ns] is consumed before the reference writes it (its malloc'd value
// was 0 in the run that produced testwaveform.dat), so it starts at 0 here.
int waveform(int ns, double st, double *t, double *dt, double *wave) {
  if (ns < 1 || !(st > 0)) return fail(TCSEM_EINVAL, "bad source discretisation");
  for (int i = 0; i < 2 * ns + 2; ++i) dt[i] = 0.0;
  t[0] = 0;
  for (int i = 0; i < ns; ++i) {
    const double t0 = t[i] - 3 * st;
    dt[i] = -st * st * (1 - std::exp(-4.5)) / ns / t0 / std::exp(-t0 * t0 / (2 * st * st));
    t[i + 1] = t[i] + dt[i];
  }
  if (3 * st < t[ns]) t[ns] = 3 * st;
  for (int i = ns + 1; i < 2 * ns + 1; ++i) {
    t[i] = 6 * st - t[2 * ns - i];
    dt[i - 1] = t[i] - t[i - 1];
  }
  for (int i = 0; i < 2 * ns + 1; ++i)
    wave[i] = 1 / (std::sqrt(2 * kPi) * st) * std::exp(-(t[i] - 3 * st) * (t[i] - 3 * st) / 2 / st / st);
  dt[2 * ns + 1] = 0.5 * (dt[2 * ns] * dt[2 * ns] / dt[2 * ns - 1] + dt[2 * ns]);
  double t0 = dt[0];
  dt[0] = 0.5 * (t0 * t0 / dt[1] + t0);
  for (int i = 1; i < 2 * ns + 1; ++i) {
    const double t1 = dt[i];
    dt[i] = 0.5 * (t0 + dt[i]);
    t0 = t1;
  }
  return TCSEM_OK;
}

int search_insert(const double *A, int n, double target) {  // logspace.cpp:30-64
  if (n < 1) return 0;
  int low = 0, high = n - 1, pos = 0;
  while (low <= high) {
    const int mid = low + (high - low) / 2;
    if (A[mid] == target) return mid;
    if (A[mid] > target) {
      high = mid - 1;
      pos = mid;
    } else {
      low = mid + 1;
      pos = mid + 1;
    }
  }
  return pos;
}

void sample_times(const tcsem_rxtime &rt, std::vector<double> &tm) {  // LoadRxTxPos.cpp:438-447
  tm.resize(rt.Nt);
  if (rt.is_log) {
    const double step = (std::log10(rt.Tn) - std::log10(rt.T0)) / (rt.Nt - 1);
    for (int i = 0; i < rt.Nt; ++i) tm[i] = std::pow(10.0, std::log10(rt.T0) + i * step);
  } else {
    const double step = (rt.Tn - rt.T0) / (rt.Nt - 1);
    for (int i = 0; i < rt.Nt; ++i) tm[i] = rt.T0 + i * step;
  }
}

// Natural cubic spline of SplineGridConvIR2R.cu (node spacing 0.5 d[i] + 0.5 d[i+1], natural
// ends, cubic inside the window, linear beyond it) evaluated at np points.  It is linear in ys,
// so the window operator is tabulated once by evaluating it on unit vectors.
void spline_eval(int n, const double *d, const double *xs, const double *ys, int np, const double *xnew,
                 const int *idraw, double *out) {
  std::vector<double> dl(n), dd(n), du(n), rhs(n), cp(n), rp(n), b(n);
  for (int i = 0; i < n; ++i) {
    if (i == 0 || i == n - 1) {
      dd[i] = 2.0;
      dl[i] = du[i] = rhs[i] = 0.0;
    } else {
      dd[i] = 2.0 / 3.0 * (d[i] + 0.5 * d[i - 1] + 0.5 * d[i + 1]);
      dl[i] = 1.0 / 3.0 * (0.5 * d[i - 1] + 0.5 * d[i]);
      du[i] = 1.0 / 3.0 * (0.5 * d[i] + 0.5 * d[i + 1]);
      rhs[i] = (ys[i + 1] - ys[i]) / (0.5 * d[i] + 0.5 * d[i + 1]) - (ys[i] - ys[i - 1]) / (0.5 * d[i - 1] + 0.5 * d[i]);
    }
  }
  cp[0] = du[0] / dd[0];
  rp[0] = rhs[0] / dd[0];
  for (int i = 1; i < n; ++i) {
    const double den = dd[i] - dl[i] * cp[i - 1];
    cp[i] = du[i] / den;
    rp[i] = (rhs[i] - dl[i] * rp[i - 1]) / den;
  }
  b[n - 1] = rp[n - 1];
  for (int i = n - 2; i >= 0; --i) b[i] = rp[i] - cp[i] * b[i + 1];
  for (int p = 0; p < np; ++p) {
    const int id = std::max(idraw[p] - 1, 0);
    double ma, mc;
    if (id < n - 1) {
      const double h = 0.5 * d[id] + 0.5 * d[id + 1];
      ma = 1.0 / 3.0 * (b[id + 1] - b[id]) / h;
      mc = (ys[id + 1] - ys[id]) / h - 1.0 / 3.0 * (2.0 * b[id] + b[id + 1]) * h;
    } else {
      const double h = 0.5 * d[id - 1] + 0.5 * d[id];
      ma = 1.0 / 3.0 * (b[id] - b[id - 1]) / h;
      mc = (ys[id] - ys[id - 1]) / h - 1.0 / 3.0 * (2.0 * b[id - 1] + b[id]) * h;
      mc += 3.0 * ma * h * h + 2.0 * b[id - 1] * h;
      ma = 0.0;
    }
    const double hh = xnew[p] - xs[id];
    const int ir = idraw[p] - 1;
    if (ir >= 0 && ir < n - 1) out[p] = ma * hh * hh * hh + b[ir] * hh * hh + mc * hh + ys[ir];
    else out[p] = (b[std::max(ir, 0)] * hh + mc) * hh + ys[std::max(ir, 0)];
  }
}

// weights W[p][q]: the window operator on unit vectors
void spline_weights(int n, const double *d, const double *xs, int np, const double *xnew, const int *idraw,
                    std::vector<double> &W) {
  W.assign((size_t)np * n, 0.0);
  std::vector<double> e(n, 0.0), col(np);
  for (int q = 0; q < n; ++q) {
    e[q] = 1.0;
    spline_eval(n, d, xs, e.data(), np, xnew, idraw, col.data());
    for (int p = 0; p < np; ++p) W[(size_t)p * n + q] = col[p];
    e[q] = 0.0;
  }
}

// host radix-2 FFT (plan building only)
void fft_pow2_host(int n, std::vector<double> &re, std::vector<double> &im, int sign) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int h = len >> 1;
    for (int k = 0; k < h; ++k) {
      const double ang = sign * 2.0 * kPi * k / len, wr = std::cos(ang), wi = std::sin(ang);
      for (int i = k; i < n; i += len) {
        const double xr = re[i + h] * wr - im[i + h] * wi, xi = re[i + h] * wi + im[i + h] * wr;
        re[i + h] = re[i] - xr;
        im[i + h] = im[i] - xi;
        re[i] += xr;
        im[i] += xi;
      }
    }
  }
}

struct HostPlan {
  int n = 0, m = 0, log2m = 0;
  std::vector<double> tw, chirp, bhat;  // interleaved complex
};

void make_plan(int n, HostPlan &P) {
  P.n = n;
  P.m = 1;
  P.log2m = 0;
  const bool pow2 = (n & (n - 1)) == 0;
  const int need = pow2 ? n : 2 * n - 1;
  while (P.m < need) {
    P.m <<= 1;
    P.log2m++;
  }
  P.tw.resize(2 * P.m);  // m complex
  for (int k = 0; k < P.m; ++k) {
    const double ang = -2.0 * kPi * k / P.m;
    P.tw[2 * k] = std::cos(ang);
    P.tw[2 * k + 1] = std::sin(ang);
  }
  if (pow2) return;
  P.chirp.resize(2 * n);
  for (int k = 0; k < n; ++k) {
    const long long q = ((long long)k * k) % (2LL * n);
    const double ang = -kPi * (double)q / n;
    P.chirp[2 * k] = std::cos(ang);
    P.chirp[2 * k + 1] = std::sin(ang);
  }
  std::vector<double> br(P.m, 0.0), bi(P.m, 0.0);
  br[0] = P.chirp[0];
  bi[0] = -P.chirp[1];
  for (int k = 1; k < n; ++k) {
    br[k] = br[P.m - k] = P.chirp[2 * k];
    bi[k] = bi[P.m - k] = -P.chirp[2 * k + 1];
  }
  fft_pow2_host(P.m, br, bi, -1);
  P.bhat.resize(2 * P.m);
  for (int k = 0; k < P.m; ++k) {
    P.bhat[2 * k] = br[k] / P.m;
    P.bhat[2 * k + 1] = bi[k] / P.m;
  }
}

// located transmitter (LoadRxTxPos.cpp:143-428)
struct TxLoc {
  int ort = 0;
  double BipoLen = 0, x = 0, y = 0, z = 0, azi = 0, dip = 0;
  int ixa = 0, iya = 0, iza = 0, ixb = 0, iyb = 0, izb = 0, ictx = 0, icty = 0, ictz = 0;
  double BipoLen1 = 0, x1 = 0, y1 = 0, z1 = 0, azi1 = 0, dip1 = 0;
  int ixa1 = 0, iya1 = 0, iza1 = 0, ixb1 = 0, iyb1 = 0, izb1 = 0, ictx1 = 0, icty1 = 0, ictz1 = 0;
  std::vector<int> rx;  // 0-based receiver indices
  std::vector<SrcCell> cells;
};

void order_pair(int &a, int &b) {  // LoadRxTxPos.cpp:297-302
  if (b > a) --b;
  else if (b < a) --a;
}

}  // namespace

// ---- context -------------------------------------------------------------------------------------
struct tcsem_ctx {
  int prec = 64;
  int dev = 0;  // the device current at tcsem_create
  size_t esz = 8;
  tcsem_setup setup{};
  HostGrid g;
  std::vector<double> cond;
  int PX = 0, PY = 0;
  long long PS = 0, fstride = 0, pstride = 0, wstride = 0;
  int by = 8, pf = 2, ntx = 0, nty = 0, nper = 0;  // march tiling
  int tune_nbatch = 0, tune_by = 0, tune_pf = 0, tune_nch = 0, tune_fuse = 1;  // autotuned tiling, its batch size
  int tune_s8 = 1;
  int stx = 0, sty = 0, sper = 0;                   // surface tiling
  int nchunk = 0, chunk_len = 0;                    // z-chunks of the lower march (0: one pass)
  int fuse = 1;  // chunked march: start chunks from the chunk sums of B (1) or a k_zsum pass (0)
  int xm = 1;                                       // x margin of the layout in elements
  int nslots = 0;
  // stream: the volume march and every host-ordered operation; stream2: upward continuation,
  // surface level and receiver gathers of a step, overlapping the lower march of the next step
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t xev[2] = {};  // [0] march of the step done (stream), [1] stream2 work of the step done
  // device buffers
  char *d_fields = nullptr;  // 2 states x 5 fields x nslots x fstride elements
  int cur = 0;               // state holding the current fields
  char *d_sig = nullptr;
  // the conductivity as an 8-bit index into a table of its distinct values (in the field
  // precision; entry 0 is 0), for models of at most 255 distinct nonzero values: the march reads
  // one byte per cell instead of a word.  TCSEM_SIGMA8=0 leaves it off.
  unsigned char *d_sig8 = nullptr;
  char *d_slut = nullptr;
  int sig8_n = 0;  // table entries in use
  bool s8 = true;  // the march reads the index (when there is one): the autotuner's choice
  // bz0, bxair, byair, carry sum, carry compensation, and the second carry sum and compensation of
  // a two-step pass: 7 x nslots x pstride
  char *d_planes = nullptr;
  char *d_zs = nullptr;      // z-chunk Bz partial sums: (kMaxChunks - 1) x nslots x pstride
  // chunk sums of B written by the chunked lower march for the next step (FP64): 2 parities x
  // (ib_cap chunks) x (x, y) x nslots x pstride; valid while the state is only advanced by steps
  // of the same chunking (set, reset, begin and model-batch steps invalidate them; the source pass
  // moves them with the B cells it recomputes)
  double *d_ib = nullptr;
  int ib_cap = 0, ib_par = 0, ib_nchunk = 0, ib_chunk_len = 0;
  bool ib_valid = false;
  char *d_rd = nullptr;      // rdx, rdx2, rdy, rdy2, dzt, rdz, rdz2 (field type)
  double *d_grid = nullptr;  // X Y Z dx dy dz Xn Yn
  int *d_maps = nullptr;     // xBx2N yo2N xo2N yBy2N, then the nyr regular rows the bilinear reads
  int nyr = 0;
  double *d_up = nullptr;    // wl wr wf wb gx gy, plan tables
  // upward-continuation work: the tiled chain's A, G, bxr, byr regions (slot strides sA, sG, sX
  // doubles), or the line chain's cplx (complex), tx, bxr, byr regions (wstride each)
  double *d_work = nullptr;
  int up_tiled = 0, MA = 0, MG = 0, XR = 0;
  long long sA = 0, sG = 0, sX = 0;
  double *d_scratch = nullptr;
  size_t scratch_bytes = 0;
  UpcontArgs<double> up64{};
  UpcontArgs<float> up32{};
  HostPlan plx, ply;
  std::vector<double> gxT, gyT;  // gains kx-major with the non-Hermitian bins zeroed (tiled chain)
  std::vector<double> wl, wr, wf, wb, gx, gy;
  int n1 = 0, n2 = 0, n3 = 0;
  // receivers
  bool have_rx = false;
  tcsem_rxtime rt{};
  int nrx = 0, ncomp = 0, nefcmp = 0, comps[6] = {0};
  std::vector<double> rxx, rxy, rxz, times;
  std::vector<int> rix, riy, riz;
  double *d_rx = nullptr;
  int *d_rxi = nullptr;
  // transmitters
  std::vector<TxLoc> tx;
  // run state
  int tx_first = 0, nbatch = 0, nrmax = 0;
  bool begun = false;
  int n = 0, steps = 0, ne = 0, nb = 0;
  double toT = 0, dt1 = 0, dt2 = 0, Tpof[2][3] = {{0}};
  std::vector<double> wt, wdt, wwave;
  double *d_ef0 = nullptr, *d_efs = nullptr;
  SrcCell *d_src = nullptr;
  int *d_srcidx = nullptr;  // first[nbatch], count[nbatch]
  // source pass after the march (levels 1..N-1): merged E cells, the B cells they touch, and
  // the (first, count) runs of those B cells, one per component and column
  SrcCell *d_fixe = nullptr, *d_fixb = nullptr;
  int2 *d_fixg = nullptr;
  int *d_fixidx = nullptr;  // E first[nbatch], E count[nbatch], run first[nbatch], run count[nbatch]
  bool has_fix = false;
  int *d_rxl = nullptr;     // rx_local, rx_first[nbatch], rx_count[nbatch]
  int *d_ops = nullptr;     // it of every op of the run (E ops then B ops)
  double *d_opl1 = nullptr;
  int src_kmin = 1, src_kmax = 0;
  std::vector<int> op_it;
  std::vector<double> op_l1;
  // timing: intervals between pooled events recorded on the streams while stepping (no extra
  // synchronisation), folded into tacc/tn when read
  bool timing = false;
  struct Span {
    int kind, count;  // kind: 0 march, 1 surface, 2 upward continuation, 3 receiver gathers, 4 two-step pass
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  std::vector<Span> spans;
  double tacc[5] = {0, 0, 0, 0, 0};
  long tn[5] = {0, 0, 0, 0, 0};
  // graph replays (tcsem_set_graph): runs of graph_k steady-state steps with no receiver sample
  // are captured once per starting state and replayed; their step scalars come from d_gs
  int graph_k = 0;
  double *d_gs = nullptr;
  // two steps per lower pass (tcsem_set_pairs): workgroup rows (0: off) and prefetch depth of
  // k_march2, its tiles, and the intermediate state (5 fields x nslots x fstride, allocated when
  // first used; only levels 0 .. 2 are written)
  int pair_by = 0, pair_pf = 1, pair_ntx = 0, pair_nty = 0, pair_nper = 0;
  int pair_set = -1, pair_set_pf = 1;  // tcsem_set_pairs' choice (-1: none), over any other
  int tune_pair_by = 0, tune_pair_pf = 1;  // the autotuner's, for its batch size tune_nbatch
  char *d_mid = nullptr;
  long pairs_done = 0;
  struct Graph {
    int cur, ib_par, ib_valid, by, pf, nchunk, chunk_len, fuse, s8;  // starting state and tiling
    int cur_after, ib_par_after, ib_valid_after;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  long graph_replays = 0;
};

namespace {

// Compute units of the current device (256 on MI355X): the tiling heuristics size their grids by it.
int device_cus() {
  int cu = 256, dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
    cu = prop.multiProcessorCount;
  return cu;
}

// Rows per march tile that the precision is built for: FP64 4, 8, 10 or 12 (16 do not fit its
// registers), FP32 4, 8, 10 or 16; other counts run as 8.
int supported_rows(int prec, int by) {
  if (prec == 64) return by == 4 || by == 10 || by == 12 ? by : 8;
  return by == 4 || by == 10 || by == 16 ? by : 8;
}

// Launch geometry of a march tiling: `by` rows per workgroup (by - 2 owned), prefetch depth pf,
// nch z-chunks of the lower launch (1: one pass).
void apply_tiling(tcsem_ctx *c, int by, int pf, int nch) {
  const int L = c->g.info.L, M = c->g.info.M, N = c->g.info.N, nlev = N - 2;
  by = supported_rows(c->prec, by);
  c->by = by;
  c->pf = pf;
  nch = std::max(1, std::min({nch, kMaxChunks, N >= 6 ? nlev : 1}));
  if (nch > 1) {
    c->chunk_len = (nlev + nch - 1) / nch;
    c->nchunk = (nlev + c->chunk_len - 1) / c->chunk_len;  // no empty chunk
  } else {
    c->nchunk = 0;
    c->chunk_len = 0;
  }
  if (c->nchunk == 1) c->nchunk = 0;
  // march: 64-column blocks over columns 0..L-1 and rows 0..M-1 (column L and row M hold no
  // updated component); surface: 62 owned columns over 0..L, rows 0..M
  c->ntx = (L + kMarchLanes - 1) / kMarchLanes;
  c->nty = (M + (by - 2) - 1) / (by - 2);
  c->nper = (c->ntx * c->nty + 7) / 8;
  c->stx = (L + 1 + kSurfaceOwnX - 1) / kSurfaceOwnX;
  c->sty = (M + 1 + (kSurfaceRows - 2) - 1) / (kSurfaceRows - 2);
  c->sper = (c->stx * c->sty + 7) / 8;
}

// Tiles of the two-step pass (rows 0: pairs off).
void pair_tiling(tcsem_ctx *c, int by, int pf) {
  const int L = c->g.info.L, M = c->g.info.M;
  c->pair_by = by;
  c->pair_pf = pf;
  if (!by) return;
  c->pair_ntx = (L + kMarchLanes - 5) / (kMarchLanes - 4);
  c->pair_nty = (M + by - 5) / (by - 4);
  c->pair_nper = (c->pair_ntx * c->pair_nty + 7) / 8;
}

// Tiling of an nbatch-transmitter step: the autotuned one for this batch size when there is
// one, else the heuristic (10-row tiles for FP32 and 12-row tiles for FP64 at prefetch 1, the
// autotuner's pick on every single model and transmitter batch measured; 4-row tiles at prefetch
// 2 for FP32 steps under 2^21 cell-updates, whose march is launch-latency bound; z-chunks when the
// tiles alone leave CUs idle); the TCSEM_MARCH_ROWS / _PF / _ZCHUNKS environment variables
// override either.
void choose_tiling(tcsem_ctx *c, int nbatch) {
  const int L = c->g.info.L, M = c->g.info.M, N = c->g.info.N, nlev = N - 2;
  const int cu = device_cus();
  auto tiles = [&](int b) { return (long long)((L + kMarchLanes - 1) / kMarchLanes) * ((M + (b - 2) - 1) / (b - 2)); };
  int by = c->prec == 64 ? 12 : 10, pf = 1;
  if (c->prec == 32 && (long long)L * M * N * nbatch < (1LL << 21)) {
    by = 4;
    pf = 2;
  }
  // enough chunks for two workgroups per CU, each chunk at least 8 levels deep
  int nch = 1;
  const long long wg = tiles(by) * nbatch;
  if (N >= 6 && wg < 2LL * cu) nch = (int)std::min<long long>((2LL * cu + wg - 1) / wg, std::max(1, nlev / 8));
  int fuse = 1;
  c->s8 = true;
  if (c->tune_nbatch == nbatch) {
    by = c->tune_by;
    pf = c->tune_pf;
    nch = c->tune_nch;
    fuse = c->tune_fuse;
    c->s8 = c->tune_s8 != 0;
  }
  if (const char *e = std::getenv("TCSEM_MARCH_ZSUM")) {  // "pass": k_zsum pass, "fused": chunk sums
    if (!std::strcmp(e, "pass")) fuse = 0;
    if (!std::strcmp(e, "fused")) fuse = 1;
  }
  c->fuse = fuse;
  if (const char *e = std::getenv("TCSEM_MARCH_ROWS")) {
    const int v = std::atoi(e);
    if (v == 4 || v == 8 || v == 10 || v == 12 || v == 16) by = v;
    if (supported_rows(c->prec, by) != by)
      std::fprintf(stderr, "tcsem: TCSEM_MARCH_ROWS=%d has no FP%d march; running %d rows\n", by, c->prec,
                   supported_rows(c->prec, by));
  }
  if (const char *e = std::getenv("TCSEM_MARCH_PF")) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 4) pf = v;
  }
  if (const char *e = std::getenv("TCSEM_MARCH_ZCHUNKS")) {
    const int v = std::atoi(e);
    if (v >= 1) nch = v;
  }
  apply_tiling(c, by, pf, nch);
  // Two steps per pass: tcsem_set_pairs' choice, else TCSEM_MARCH_PAIRS ("0", or "rows,pf"), else
  // the autotuner's for this batch size, else 12-row passes at prefetch 1 when they give at least
  // two workgroups per CU (faster than one step at a time on every such workload measured; fewer
  // leave CUs idle, and the one-step march's z-chunks serve those grids better). The pass holds
  // one workgroup per CU and a round of them takes about as long as the tile has rows, so 14-row
  // (FP32) or 11-row (FP64) passes replace the 12-row ones where their rounds x rows is 5% or
  // more smaller: a plane whose 12-row tiles leave the last round mostly empty (512 x 512).
  int pby = 0, ppf = 1;
  auto ptiles = [&](int b) {
    return (long long)((L + kMarchLanes - 5) / (kMarchLanes - 4)) * ((M + b - 5) / (b - 4)) * nbatch;
  };
  if (ptiles(12) >= 2LL * cu) {
    const int alt = c->prec == 32 ? 14 : 11;
    auto cost = [&](int b) { return (double)((ptiles(b) + cu - 1) / cu) * b; };
    pby = cost(alt) <= 0.95 * cost(12) ? alt : 12;
  }
  if (c->tune_nbatch == nbatch) {
    pby = c->tune_pair_by;
    ppf = c->tune_pair_pf;
  }
  if (const char *e = std::getenv("TCSEM_MARCH_PAIRS")) {
    int r = 0, q = 1;
    if (std::sscanf(e, "%d,%d", &r, &q) >= 1 && (r == 0 || (pair_rows_supported(c->prec, r) && q >= 1 && q <= 2))) {
      pby = r;
      ppf = q;
    } else {
      std::fprintf(stderr, "tcsem: TCSEM_MARCH_PAIRS=%s is not 0 or rows,pf (rows 8, 10, 12, 14, 16 (FP64: 8, 10, 11, 12); pf 1, 2)\n", e);
    }
  }
  if (c->pair_set >= 0) {
    pby = c->pair_set;
    ppf = c->pair_set_pf;
  }
  if (N < 6) pby = 0;
  pair_tiling(c, pby, ppf);
}

template <typename T>
MarchArgs<T> march_args(tcsem_ctx *c) {
  MarchArgs<T> a{};
  const long long org = c->PX + c->xm;
  const long long fs = c->fstride * c->nslots;
  T *fi = (T *)c->d_fields + (long long)c->cur * 5 * fs;
  T *fo = (T *)c->d_fields + (long long)(1 - c->cur) * 5 * fs;
  a.exi = fi + org;
  a.eyi = fi + fs + org;
  a.ezi = fi + 2 * fs + org;
  a.bxi = fi + 3 * fs + org;
  a.byi = fi + 4 * fs + org;
  a.ex = fo + org;
  a.ey = fo + fs + org;
  a.ez = fo + 2 * fs + org;
  a.bx = fo + 3 * fs + org;
  a.by = fo + 4 * fs + org;
  a.sig = (const T *)c->d_sig + org;
  a.sig8 = c->d_sig8 && c->s8 ? c->d_sig8 + org : nullptr;
  a.slut = (const T *)c->d_slut;
  T *pl = (T *)c->d_planes;
  const long long ps = c->pstride * c->nslots;
  a.bz0 = pl + org;
  a.bxair = pl + ps + org;
  a.byair = pl + 2 * ps + org;
  a.cs = pl + 3 * ps + org;
  a.cc = pl + 4 * ps + org;
  a.zs = (T *)c->d_zs + org;
  a.zstride = ps;
  a.nchunk = 0;
  a.chunk_len = 0;
  a.k_hi = c->g.info.N - 1;
  a.k_lo = 1;
  a.carry = 0;
  a.fstride = c->fstride;
  a.pstride = c->pstride;
  a.PS = c->PS;
  a.L = c->g.info.L;
  a.M = c->g.info.M;
  a.N = c->g.info.N;
  a.PX = c->PX;
  a.xm = c->xm;
  a.rows = c->by;
  a.pf = c->pf;
  a.ntx = c->ntx;
  a.nty = c->nty;
  a.nper = c->nper;
  a.stx = c->stx;
  a.sty = c->sty;
  a.sper = c->sper;
  const T *rd = (const T *)c->d_rd;
  const int L = a.L, M = a.M, N = a.N;
  a.rdx = rd;
  a.rdx2 = rd + (L + 3);
  a.rdy = rd + 2 * (L + 3);
  a.rdy2 = rd + 2 * (L + 3) + (M + 3);
  a.dzt = rd + 2 * (L + 3) + 2 * (M + 3);
  a.rdz = a.dzt + N;
  a.rdz2 = a.dzt + 2 * N;
  a.src = c->d_src;
  a.src_first = c->d_srcidx;
  a.src_count = c->d_srcidx ? c->d_srcidx + std::max(c->nbatch, 1) : nullptr;
  a.src_kmin = c->src_kmin;
  a.src_kmax = c->src_kmax;
  return a;
}

template <typename T>
UpcontArgs<T> upcont_args(tcsem_ctx *c) {
  UpcontArgs<T> u{};
  const long long org = c->PX + c->xm;
  T *pl = (T *)c->d_planes;
  const long long ps = c->pstride * c->nslots;
  u.bz0 = pl + org;
  u.bxair = pl + ps + org;
  u.byair = pl + 2 * ps + org;
  u.pstride = c->pstride;
  const tcsem_grid_info &I = c->g.info;
  u.L = I.L;
  u.M = I.M;
  u.PX = c->PX;
  u.pz = u.pxa = u.pya = c->PX;
  u.L1 = I.L1;
  u.M1 = I.M1;
  u.AnLx = I.AnLx_New;
  u.AnRx = I.AnRx_New;
  u.Any = I.Any_New;
  u.AL = c->g.AL;
  u.AY = c->g.AY;
  u.Nx = c->setup.Nx;
  u.Ny = c->setup.Ny;
  u.n1 = c->n1;
  u.n2 = c->n2;
  u.n3 = c->n3;
  double *p = c->d_up;
  u.wl = p;
  p += c->wl.size();
  u.wr = p;
  p += c->wr.size();
  u.wf = p;
  p += c->wf.size();
  u.wb = p;
  p += c->wb.size();
  u.gx = p;
  p += c->gx.size();
  u.gy = p;
  p += c->gy.size();
  auto plan = [&](const HostPlan &H, FftPlan &P) {
    P.n = H.n;
    P.m = H.m;
    P.log2m = H.log2m;
    P.tw = p;
    p += H.tw.size();
    P.chirp = H.chirp.empty() ? nullptr : p;
    p += H.chirp.size();
    P.bhat = H.bhat.empty() ? nullptr : p;
    p += H.bhat.size();
  };
  plan(c->plx, u.px);
  plan(c->ply, u.py);
  u.H = I.L1 / 2 + 1;
  u.wstride = c->wstride;
  u.gxT = p;
  p += c->gxT.size();
  u.gyT = p;
  p += c->gyT.size();
  double *w = c->d_work;
  u.tiled = c->up_tiled;
  if (c->up_tiled) {
    const long long ns = c->nslots;
    T *wt = (T *)w;  // tiled planes in the field type; complex entries are T pairs
    u.A = wt;
    u.sA = c->sA;
    u.MA = c->MA;
    u.G = wt + 2 * ns * c->sA;
    u.sG = c->sG;
    u.MG = c->MG;
    u.tbx = wt + 2 * ns * (c->sA + c->sG);
    u.tby = u.tbx + ns * c->sX;
    u.sX = c->sX;
    u.XR = c->XR;
  } else {
    const long long ws = c->wstride * c->nslots;
    u.cplx = (double2 *)w;
    u.tx = w + 2 * ws;
    u.bxr = w + 3 * ws;
    u.ry = u.bxr;  // consumed by the forward column FFTs before the results are written
    u.byr = w + 4 * ws;
    u.sX = c->wstride;
    u.XR = I.L1;
  }
  const double *gd = c->d_grid;
  u.X = gd;
  u.Y = gd + I.L;
  u.dx = gd + I.L + I.M + I.N;
  u.dy = u.dx + I.L;
  u.Xn = u.dy + I.M + I.N;
  u.Yn = u.Xn + I.L1;
  const int *mp = c->d_maps;
  u.xBx2N = mp;
  u.yo2N = mp + (I.L + 1);
  u.xo2N = u.yo2N + I.M;
  u.yBy2N = u.xo2N + I.L;
  u.yrows = u.yBy2N + I.M + 1;
  u.nyr = c->nyr;
  return u;
}

// Graph executables bake device addresses: dropped whenever an allocation they use may change.
void drop_graphs(tcsem_ctx *c) {
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

void free_run(tcsem_ctx *c) {
  drop_graphs(c);
  (void)hipFree(c->d_ef0);
  (void)hipFree(c->d_efs);
  (void)hipFree(c->d_src);
  (void)hipFree(c->d_srcidx);
  (void)hipFree(c->d_fixe);
  (void)hipFree(c->d_fixb);
  (void)hipFree(c->d_fixg);
  (void)hipFree(c->d_fixidx);
  c->d_fixe = c->d_fixb = nullptr;
  c->d_fixg = nullptr;
  c->d_fixidx = nullptr;
  c->has_fix = false;
  (void)hipFree(c->d_rxl);
  (void)hipFree(c->d_ops);
  (void)hipFree(c->d_opl1);
  c->d_ef0 = c->d_efs = nullptr;
  c->d_src = nullptr;
  c->d_srcidx = c->d_rxl = c->d_ops = nullptr;
  c->d_opl1 = nullptr;
  c->begun = false;
}

int alloc_slots(tcsem_ctx *c, int nslots) {
  if (nslots <= c->nslots) return TCSEM_OK;
  drop_graphs(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(c->d_fields);
  (void)hipFree(c->d_planes);
  (void)hipFree(c->d_work);
  (void)hipFree(c->d_zs);
  (void)hipFree(c->d_ib);
  (void)hipFree(c->d_mid);
  c->d_mid = nullptr;
  c->d_ib = nullptr;
  c->ib_cap = 0;
  c->ib_valid = false;
  c->d_fields = c->d_planes = c->d_zs = nullptr;
  c->d_work = nullptr;
  c->nslots = 0;
  const size_t fb = (size_t)2 * 5 * nslots * c->fstride * c->esz;
  const size_t pb = (size_t)7 * nslots * c->pstride * c->esz;
  // bytes per slot: tiled chain A and G (complex of the field type) and Bx, By (field type);
  // line chain cplx (complex FP64), tx, bxr (also the y-padding rows), byr (FP64)
  const size_t per_slot = c->up_tiled ? (size_t)(2 * c->sA + 2 * c->sG + 2 * c->sX) * c->esz
                                      : (size_t)5 * c->wstride * sizeof(double);
  const size_t wb = (size_t)nslots * per_slot;
  if (hipMalloc(&c->d_fields, fb) != hipSuccess) return fail(TCSEM_ENOMEM, "field allocation of %zu bytes failed", fb);
  if (hipMalloc(&c->d_planes, pb) != hipSuccess) return fail(TCSEM_ENOMEM, "plane allocation failed");
  if (hipMalloc((void **)&c->d_work, wb) != hipSuccess) return fail(TCSEM_ENOMEM, "work allocation failed");
  const size_t zb = (size_t)(kMaxChunks - 1) * nslots * c->pstride * c->esz;
  if (hipMalloc(&c->d_zs, zb) != hipSuccess) return fail(TCSEM_ENOMEM, "chunk-sum allocation failed");
  HIPCHK(launch_fill_zero(c->d_zs, zb, c->stream));
  HIPCHK(launch_fill_zero(c->d_fields, fb, c->stream));
  HIPCHK(launch_fill_zero(c->d_planes, pb, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->nslots = nslots;
  c->cur = 0;
  return TCSEM_OK;
}

template <typename T>
void put(std::vector<char> &buf, size_t idx, double v) {
  ((T *)buf.data())[idx] = (T)v;
}

int upload_double_as(tcsem_ctx *c, char *dst, const std::vector<double> &v) {
  if (c->prec == 64) {
    HIPCHK(hipMemcpy(dst, v.data(), v.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(v.begin(), v.end());
    HIPCHK(hipMemcpy(dst, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  return TCSEM_OK;
}

// source cells of a located transmitter: AddTx_x/y/z (ElectricalFiled.cu:347-475) with the
// currents of the host switch (ElectricalFiled.cu:532-641).  Per step the cell receives
// E -= Wave[n] * w.  The second bipole of ort 5 is injected on the first bipole's cells, as the
// reference's kernels read only the first bipole's indices.
void source_cells(const tcsem_ctx *c, TxLoc &t) {
  const HostGrid &g = c->g;
  const int L = g.info.L, M = g.info.M;
  const double *cd = c->cond.data();
  auto C = [&](int k, int j, int i) { return cd[((size_t)k * M + j) * L + i]; };
  t.cells.clear();
  auto addx = [&](double cur, int nthreads) {
    const int ixa = std::min(t.ixa, t.ixb), iya = t.icty, iza = t.ictz;
    for (int ix = 0; ix < nthreads; ++ix) {
      if (ix > std::abs(t.ixb - t.ixa)) continue;
      double s = (iza == 0) ? C(iza, iya, ix + ixa) + C(iza, iya - 1, ix + ixa)
                            : C(iza, iya, ix + ixa) + C(iza, iya - 1, ix + ixa) + C(iza - 1, iya, ix + ixa) +
                                  C(iza - 1, iya - 1, ix + ixa);
      t.cells.push_back({0, iza, iya, ix + ixa, 2.0 * cur / (0.25 * s * g.dy[iya] * g.dz[iza])});
    }
  };
  auto addy = [&](double cur, int nthreads) {
    const int ixa = t.ictx, iya = std::min(t.iya, t.iyb), iza = t.ictz;
    for (int iy = 0; iy < nthreads; ++iy) {
      if (iy > std::abs(t.iyb - t.iya)) continue;
      double s = (iza == 0) ? C(iza, iy + iya, ixa - 1) + C(iza, iy + iya, ixa)
                            : C(iza, iy + iya, ixa - 1) + C(iza, iy + iya, ixa) + C(iza - 1, iy + iya, ixa - 1) +
                                  C(iza - 1, iy + iya, ixa);
      t.cells.push_back({1, iza, iy + iya, ixa, 2.0 * cur / (0.25 * s * g.dx[ixa] * g.dz[iza])});
    }
  };
  auto addz = [&](double cur, int nthreads) {
    const int ixa = t.ictx, iya = t.icty, iza = std::min(t.iza, t.izb);
    for (int iz = 0; iz < nthreads; ++iz) {
      if (iz > std::abs(t.izb - t.iza)) continue;
      double s = C(iza + iz, iya, ixa - 1) + C(iza + iz, iya, ixa) + C(iza + iz, iya - 1, ixa - 1) +
                 C(iza + iz, iya - 1, ixa);
      t.cells.push_back({2, iza + iz, iya, ixa, 2.0 * cur / (0.25 * s * g.dx[ixa] * g.dy[iya])});
    }
  };
  const double *X = g.X.data(), *Y = g.Y.data(), *Z = g.Z.data(), *dx = g.dx.data(), *dy = g.dy.data(),
               *dz = g.dz.data();
  switch (t.ort) {
    case 1:
      addx(1.0 / (X[t.ixb] - X[t.ixa] + dx[t.ixb] / 2 + dx[t.ixa] / 2), std::abs(t.ixb - t.ixa) + 1);
      break;
    case 2:
      addy(1.0 / (Y[t.iyb] - Y[t.iya] + dy[t.iyb] / 2 + dy[t.iya] / 2), std::abs(t.iyb - t.iya) + 1);
      break;
    case 3:
      addz(1.0 / (Z[t.izb] - Z[t.iza] + dz[t.izb] / 2 + dz[t.iza] / 2), std::abs(t.izb - t.iza) + 1);
      break;
    case 4:
    case 5:
      for (int b = 0; b < (t.ort == 5 ? 2 : 1); ++b) {
        const double dip = b ? t.dip1 : t.dip, azi = b ? t.azi1 : t.azi;
        const int ixa = b ? t.ixa1 : t.ixa, ixb = b ? t.ixb1 : t.ixb, iya = b ? t.iya1 : t.iya,
                  iyb = b ? t.iyb1 : t.iyb, iza = b ? t.iza1 : t.iza, izb = b ? t.izb1 : t.izb;
        const double f = t.ort == 5 ? (b ? t.BipoLen1 : t.BipoLen) : 1.0;
        addx(std::cos(dip) * std::cos(azi) * f / (std::fabs(X[ixb] - X[ixa]) + dx[ixb] / 2.0 + dx[ixa] / 2.0),
             std::abs(ixb - ixa) + 1);
        addy(std::cos(dip) * std::sin(azi) * f / (std::fabs(Y[iyb] - Y[iya]) + dy[iyb] / 2.0 + dy[iya] / 2.0),
             std::abs(iyb - iya) + 1);
        addz(std::sin(dip) * f / (std::fabs(Z[izb] - Z[iza]) + dz[izb] / 2.0 + dz[iza] / 2.0),
             std::abs(izb - iza) + 1);
      }
      break;
  }
  // cells with zero weight (a zero current component) are dropped
  t.cells.erase(std::remove_if(t.cells.begin(), t.cells.end(), [](const SrcCell &s) { return s.w == 0.0; }),
                t.cells.end());
}

int locate_tx(const tcsem_ctx *c, const tcsem_tx &in, TxLoc &t) {  // LoadRxTxPos.cpp:143-428
  const HostGrid &g = c->g;
  const int L = g.info.L, M = g.info.M, K = g.info.N;
  const double *X = g.X.data(), *Y = g.Y.data(), *Z = g.Z.data();
  t = TxLoc();
  t.ort = in.ort;
  t.BipoLen = in.bipolen;
  t.x = in.x;
  t.y = in.y;
  t.z = in.z;
  t.BipoLen1 = in.bipolen1;
  t.x1 = in.x1;
  t.y1 = in.y1;
  t.z1 = in.z1;
  if (in.ort == 4 || in.ort == 5) {
    t.azi = in.azi_deg * kPi / 180.0;
    t.dip = in.dip_deg * kPi / 180.0;
  }
  if (in.ort == 5) {
    t.azi1 = in.azi1_deg * kPi / 180.0;
    t.dip1 = in.dip1_deg * kPi / 180.0;
  }
  switch (in.ort) {
    case 1:
      t.ixa = search_insert(X, L, t.x - t.BipoLen / 2);
      t.ixb = search_insert(X, L, t.x + t.BipoLen / 2) - 1;
      if (t.ixa == 0 || t.ixb == L) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in x", in.id);
      t.icty = search_insert(Y, M, t.y);
      if (t.icty == 0 || t.icty == M) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in y", in.id);
      t.ictz = search_insert(Z, K, t.z);
      break;
    case 2:
      t.ictx = search_insert(X, L, t.x);
      if (t.ictx == 0 || t.ictx == L) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in x", in.id);
      t.iya = search_insert(Y, M, t.y - t.BipoLen / 2);
      t.iyb = search_insert(Y, M, t.y + t.BipoLen / 2) - 1;
      if (t.iya == 0 || t.iyb == M) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in y", in.id);
      t.ictz = search_insert(Z, K, t.z);
      break;
    case 3:
      t.ictx = search_insert(X, L, t.x);
      if (t.ictx == 0 || t.ictx == L) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in x", in.id);
      t.icty = search_insert(Y, M, t.y);
      if (t.icty == 0 || t.icty == M) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in y", in.id);
      t.iza = search_insert(Z, K, t.z - t.BipoLen / 2);
      t.izb = search_insert(Z, K, t.z + t.BipoLen / 2) - 1;
      break;
    case 4:
    case 5: {
      const double cx = t.BipoLen * std::cos(t.dip) * std::cos(t.azi) / 2.0;
      const double cy = t.BipoLen * std::cos(t.dip) * std::sin(t.azi) / 2.0;
      const double cz = t.BipoLen * std::sin(t.dip) / 2.0;
      t.ixa = search_insert(X, L, t.x - cx);
      t.ixb = search_insert(X, L, t.x + cx);
      order_pair(t.ixa, t.ixb);
      t.iya = search_insert(Y, M, t.y - cy);
      t.iyb = search_insert(Y, M, t.y + cy);
      order_pair(t.iya, t.iyb);
      t.iza = search_insert(Z, K, t.z - cz);
      t.izb = search_insert(Z, K, t.z + cz);
      order_pair(t.iza, t.izb);
      t.ictx = search_insert(X, L, t.x);
      t.icty = search_insert(Y, M, t.y);
      t.ictz = search_insert(Z, K, t.z);
      if (in.ort == 5) {
        const double ax = t.BipoLen1 * std::cos(t.dip1) * std::cos(t.azi1) / 2.0;
        const double ay = t.BipoLen1 * std::cos(t.dip1) * std::sin(t.azi1) / 2.0;
        const double az = t.BipoLen1 * std::sin(t.dip1) / 2.0;
        t.ixa1 = search_insert(X, L, t.x1 - ax);
        t.ixb1 = search_insert(X, L, t.x1 + ax);
        order_pair(t.ixa1, t.ixb1);
        t.iya1 = search_insert(Y, M, t.y1 - ay);
        t.iyb1 = search_insert(Y, M, t.y1 + ay);
        order_pair(t.iya1, t.iyb1);
        t.iza1 = search_insert(Z, K, t.z1 - az);
        t.izb1 = search_insert(Z, K, t.z1 + az);
        order_pair(t.iza1, t.izb1);
        t.ictx1 = search_insert(X, L, t.x1);
        t.icty1 = search_insert(Y, M, t.y1);
        t.ictz1 = search_insert(Z, K, t.z1);
        if (t.ixa1 == 0 || t.ixb1 == L || t.iya1 == 0 || t.iyb1 == M)
          return fail(TCSEM_EDOMAIN, "transmitter %d (second bipole) outside the model", in.id);
      }
      if (t.ixa == 0 || t.ixb == L) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in x", in.id);
      if (t.iya == 0 || t.iyb == M) return fail(TCSEM_EDOMAIN, "transmitter %d outside the model in y", in.id);
      break;
    }
    default:
      return fail(TCSEM_EINVAL, "transmitter %d: illegal orientation %d", in.id, in.ort);
  }
  // the injected cells must be interior cells of the model (the reference reads and writes
  // out of bounds otherwise)
  auto okx = [&](int i) { return i >= 1 && i < L; };
  auto oky = [&](int j) { return j >= 1 && j < M; };
  auto okz = [&](int k) { return k >= 0 && k < K; };
  bool ok = true;
  if (in.ort == 1) ok = okx(std::min(t.ixa, t.ixb)) && std::max(t.ixa, t.ixb) < L && oky(t.icty) && okz(t.ictz);
  if (in.ort == 2) ok = okx(t.ictx) && oky(std::min(t.iya, t.iyb)) && std::max(t.iya, t.iyb) < M && okz(t.ictz);
  if (in.ort == 3) ok = okx(t.ictx) && oky(t.icty) && okz(std::min(t.iza, t.izb)) && std::max(t.iza, t.izb) < K;
  if (in.ort >= 4)
    ok = okx(std::min(t.ixa, t.ixb)) && std::max(t.ixa, t.ixb) < L && oky(std::min(t.iya, t.iyb)) &&
         std::max(t.iya, t.iyb) < M && okz(std::min(t.iza, t.izb)) && std::max(t.iza, t.izb) < K &&
         okx(t.ictx) && oky(t.icty) && okz(t.ictz);
  if (!ok) return fail(TCSEM_EDOMAIN, "transmitter %d cells leave the model interior", in.id);
  t.rx.clear();
  for (int r = 0; r < in.nr; ++r) {
    const int id = in.rx_ids[r];
    if (id < 1 || id > c->nrx) return fail(TCSEM_EINVAL, "transmitter %d references receiver %d", in.id, id);
    t.rx.push_back(id - 1);
  }
  return TCSEM_OK;
}

// time-interpolation bookkeeping of EFInterp2Rx.cu:472-585 for one field group
struct Event {
  int slot_store;  // 0 or 1
  int op_first, op_count;
};

bool sample_group(tcsem_ctx *c, int fld, double t2, std::vector<int> &it, std::vector<double> &l1,
                  Event &ev) {
  int &cnt = fld == 0 ? c->ne : c->nb;
  double *T = c->Tpof[fld];
  const int Nt = c->rt.Nt;
  const std::vector<double> &tm = c->times;
  bool fire = false;
  T[2] = t2;
  while (cnt < Nt && tm[cnt] <= T[0]) ++cnt;
  if (cnt < Nt && tm[cnt] <= T[2] && tm[cnt] > T[1]) {
    ev.slot_store = 0;
    ev.op_first = (int)it.size();
    ev.op_count = 0;
    fire = true;
  } else if (cnt < Nt && tm[cnt] <= T[1] && tm[cnt] > T[0]) {
    ev.slot_store = 1;
    ev.op_first = (int)it.size();
    ev.op_count = 0;
    while (cnt < Nt && tm[cnt] <= T[1]) {
      it.push_back(cnt);
      l1.push_back((tm[cnt] - T[0]) / (T[1] - T[0]));
      ++ev.op_count;
      ++cnt;
    }
    fire = true;
  }
  T[0] = T[1];
  T[1] = T[2];
  return fire;
}

template <typename T>
UpcontArgs<T> &up_of(tcsem_ctx *c);
template <>
UpcontArgs<double> &up_of<double>(tcsem_ctx *c) {
  return c->up64;
}
template <>
UpcontArgs<float> &up_of<float>(tcsem_ctx *c) {
  return c->up32;
}

constexpr int kTimingStride = 8;

hipEvent_t timing_event(tcsem_ctx *c) {
  if (c->evused == c->evpool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->evpool.push_back(e);
  }
  return c->evpool[c->evused++];
}

// Waits for the recorded intervals and adds them to the per-kind totals.
int rc_collect(tcsem_ctx *c) {
  if (c->spans.empty()) return TCSEM_OK;
  HIPCHK(hipStreamSynchronize(c->stream2));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (const auto &sp : c->spans) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, sp.a, sp.b));
    c->tacc[sp.kind] += ms;
    c->tn[sp.kind] += sp.count;
  }
  c->spans.clear();
  c->evused = 0;
  return TCSEM_OK;
}

// The volume march of one step on stream st: with N >= 6 a lower launch (levels N-1 .. 2, z-chunked
// when the tiling says so) and, after `wait` (the previous step's stream2 work; may be null), the
// top launch (level 1 and Bz_old(0)); with fewer levels one launch.  te (null, or 4 events):
// start/end of the lower and of the top launch.
// Chunk-sum planes for nchunk chunks (allocated on first use, grown when the chunking grows; a
// new allocation is zeroed, so the plane margins and wall columns read as zero).
hipError_t ensure_ib(tcsem_ctx *c, int nchunk) {
  if (nchunk - 1 <= c->ib_cap) return hipSuccess;
  hipError_t e;
  if ((e = hipStreamSynchronize(c->stream2)) != hipSuccess || (e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return e;
  drop_graphs(c);
  (void)hipFree(c->d_ib);
  c->d_ib = nullptr;
  c->ib_cap = 0;
  c->ib_valid = false;
  const size_t bytes = (size_t)2 * (nchunk - 1) * 2 * c->nslots * c->pstride * sizeof(double);
  if ((e = hipMalloc((void **)&c->d_ib, bytes)) != hipSuccess) return e;
  if ((e = launch_fill_zero(c->d_ib, bytes, c->stream)) != hipSuccess) return e;
  c->ib_cap = nchunk - 1;
  return hipSuccess;
}

// Origin of the chunk-sum planes of parity par (chunk 0, x sums, slot 0).
double *ib_planes(tcsem_ctx *c, int par) {
  const long long ibstride = (long long)c->nslots * c->pstride;
  return c->d_ib + (long long)par * c->ib_cap * 2 * ibstride + c->PX + c->xm;
}

template <typename T>
hipError_t enqueue_march(tcsem_ctx *c, const MarchArgs<T> &a, int nb, hipStream_t st, hipEvent_t wait,
                         const hipEvent_t *te) {
  static const hipEvent_t none[4] = {};
  if (!te) te = none;
  const bool split = c->g.info.N >= 6;
  hipError_t e;
  if (split) {
    MarchArgs<T> lo = a;
    lo.k_lo = 2;
    lo.carry = 2;
    const bool chunked = c->nchunk > 1;
    bool fused = false, pass = false;
    if (chunked) {
      lo.nchunk = c->nchunk;
      lo.chunk_len = c->chunk_len;
      if (c->fuse && (e = ensure_ib(c, c->nchunk)) != hipSuccess) return e;
      // this step writes the chunk sums of its result; it reads the previous step's when they
      // describe the current state with this chunking, else the k_zsum pass supplies the partials
      fused = c->fuse && c->ib_valid && c->ib_nchunk == c->nchunk && c->ib_chunk_len == c->chunk_len;
      if (c->fuse) {
        lo.ibstride = (long long)c->nslots * c->pstride;
        lo.ibw = ib_planes(c, c->ib_par ^ 1);
        lo.ibr = fused ? ib_planes(c, c->ib_par) : nullptr;
      }
      // a k_zsum pass belongs to the march: a timed step's march interval starts at its dispatch
      pass = !fused;
      if (pass && (e = launch_zsum<T>(lo, nb, st, te[0])) != hipSuccess) return e;
      c->ib_par ^= 1;
      c->ib_valid = c->fuse != 0;
      c->ib_nchunk = c->nchunk;
      c->ib_chunk_len = c->chunk_len;
    } else {
      c->ib_valid = false;
    }
    if ((e = launch_march<T>(lo, nb, st, pass ? nullptr : te[0], te[1])) != hipSuccess) return e;
  }
  if (wait && (e = hipStreamWaitEvent(st, wait, 0)) != hipSuccess) return e;
  MarchArgs<T> hi = a;
  if (split) {
    hi.k_hi = 1;
    hi.carry = 1;
  }
  return launch_march<T>(hi, nb, st, te[2], te[3]);
}

// Step scalars of the context's current step (main.cu:223-260).
template <typename T>
void step_scalars(const tcsem_ctx *c, MarchArgs<T> &a) {
  const tcsem_setup &s = c->setup;
  const int ns = s.n_src / 2;
  const double mindx = c->g.mindx;
  a.lam = 7.5 * s.sfa2 * c->dt2 * c->dt1 / kMu0 / mindx / mindx;  // main.cu:223
  a.dt = c->dt1;
  a.dtb = 0.5 * (c->dt1 + c->dt2);
  a.amp = c->n < 2 * ns ? c->wwave[c->n] : 0.0;
}

// Time advance after a step (main.cu:253-260).
void advance_time(tcsem_ctx *c) {
  const tcsem_setup &s = c->setup;
  const int ns = s.n_src / 2;
  const double minsita = 0.1;
  c->toT += c->dt1;
  c->dt1 = c->dt2;
  c->n++;
  if (c->n < 2 * ns) c->dt2 = c->wdt[c->n + 1];
  else c->dt2 = std::min(0.15 * s.sfa1 * std::sqrt(kMu0 * c->toT * minsita / 8.) * c->g.mindx, 1.2 * c->dt1);
  c->steps++;
}

// Receiver sampling after a step (EFInterp2Rx(t0 = toT + dt1, dt = dt2)): the gathers of the
// events that fire, on stream st2.  timed: record their intervals.
template <typename T>
int sample_rx(tcsem_ctx *c, hipStream_t st2, bool timed) {
  if (!c->have_rx) return TCSEM_OK;
  const int nb = c->nbatch;
  const double t0 = c->toT + c->dt1;
  GatherArgs gA{};
  const long long org = c->PX + c->xm;
  const long long fs = c->fstride * c->nslots;
  const T *f = (const T *)c->d_fields + (long long)c->cur * 5 * fs;
  for (int q = 0; q < 5; ++q) gA.f[q] = f + q * fs + org;
  gA.f[5] = nullptr;
  gA.fstride = c->fstride;
  gA.PS = c->PS;
  gA.L = c->g.info.L;
  gA.M = c->g.info.M;
  gA.N = c->g.info.N;
  gA.PX = c->PX;
  const tcsem_grid_info &I = c->g.info;
  gA.X = c->d_grid;
  gA.Y = c->d_grid + I.L;
  gA.Z = c->d_grid + I.L + I.M;
  gA.dx = c->d_grid + I.L + I.M + I.N;
  gA.dy = gA.dx + I.L;
  gA.dz = gA.dy + I.M;
  gA.rx_first = c->d_rxl + 0;
  gA.rx_count = c->d_rxl + nb;
  gA.rx_local = c->d_rxl + 2 * nb;
  gA.rxx = c->d_rx;
  gA.rxy = c->d_rx + c->nrx;
  gA.rxz = c->d_rx + 2 * c->nrx;
  gA.rix = c->d_rxi;
  gA.riy = c->d_rxi + c->nrx;
  gA.riz = c->d_rxi + 2 * c->nrx;
  gA.ncomp = c->ncomp;
  gA.nt = c->rt.Nt;
  gA.ef0 = c->d_ef0;
  gA.efs = c->d_efs;
  gA.nrmax = c->nrmax;
  const bool has_b = c->rt.is_comp[3] || c->rt.is_comp[4] || c->rt.is_comp[5];
  for (int fld = 0; fld < 2; ++fld) {
    if (fld == 0 && c->ne >= c->rt.Nt) continue;
    if (fld == 1 && (c->nb >= c->rt.Nt || !has_b)) continue;
    Event ev{};
    const size_t before = c->op_it.size();
    if (!sample_group(c, fld, fld == 0 ? t0 : t0 + 0.5 * c->dt2, c->op_it, c->op_l1, ev)) continue;
    // upload the new ops (few per event)
    if (ev.op_count > 0) {
      HIPCHK(hipMemcpyAsync(c->d_ops + before, c->op_it.data() + before, ev.op_count * sizeof(int),
                            hipMemcpyHostToDevice, st2));
      HIPCHK(hipMemcpyAsync(c->d_opl1 + before, c->op_l1.data() + before, ev.op_count * sizeof(double),
                            hipMemcpyHostToDevice, st2));
    }
    const int c0 = fld == 0 ? 0 : c->nefcmp, c1 = fld == 0 ? c->nefcmp : c->ncomp;
    if (c1 <= c0) continue;
    gA.ncomp_group = c1 - c0;
    gA.comp_base = c0;
    for (int q = 0; q < gA.ncomp_group; ++q) gA.comps[q] = c->comps[c0 + q];
    gA.slot_store = ev.slot_store;
    gA.op_it = c->d_ops + ev.op_first;
    gA.op_l1 = c->d_opl1 + ev.op_first;
    gA.nops = ev.op_count;
    hipEvent_t g0 = nullptr, g1 = nullptr;
    if (timed) {
      g0 = timing_event(c);
      g1 = timing_event(c);
      if (!g0 || !g1) return fail(TCSEM_EDEVICE, "timing event creation failed");
      c->spans.push_back({3, 1, g0, g1});
    }
    HIPCHK(launch_gather<T>(gA, fld, nb, march_args<T>(c), st2, g0, g1));
  }
  return TCSEM_OK;
}

template <typename T>
int do_step(tcsem_ctx *c) {
  MarchArgs<T> a = march_args<T>(c);
  step_scalars<T>(c, a);
  const UpcontArgs<T> &u = up_of<T>(c);
  const int nb = c->nbatch, N = c->g.info.N;
  hipStream_t st = c->stream, st2 = c->stream2;
  // stream: lower march (levels N-1..2: reads the old state at levels >= 1 only, so it runs
  // beside the previous step's stream2 work), then, once that work is done (it writes level 0 of
  // the old state and still reads the Bz(0) plane), the top march (level 1 and Bz_old(0)).
  // stream2: upward continuation, surface level and receiver gathers of this step.
  const bool split = N >= 6;
  // timing events are recorded by the dispatch packets of the kernels themselves, on every
  // kTimingStride-th step (each timed step costs a few microseconds of queue time)
  const bool timed = c->timing && c->steps % kTimingStride == 0;
  hipEvent_t te[8] = {};
  if (timed) {
    if (c->spans.size() > 50000 && (rc_collect(c))) return TCSEM_EDEVICE;
    for (auto &e : te)
      if (!(e = timing_event(c))) return fail(TCSEM_EDEVICE, "timing event creation failed");
  }
  HIPCHK(enqueue_march<T>(c, a, nb, st, c->xev[1], te));
  // source cells at levels 1 .. N-1: a pass after the march, which moves the chunk sums of B that a
  // chunked lower launch has just written along with the B cells it recomputes
  if (c->has_fix && a.amp != 0.0) {
    MarchArgs<T> sp = a;
    sp.fixe = c->d_fixe;
    sp.fixb = c->d_fixb;
    sp.fixg = c->d_fixg;
    sp.fixidx = c->d_fixidx;
    sp.fixn = nb;
    if (split && c->nchunk > 1 && c->fuse) {
      sp.ibw = ib_planes(c, c->ib_par);  // the parity this step's lower launch wrote
      sp.ibstride = (long long)c->nslots * c->pstride;
      sp.nchunk = c->nchunk;
      sp.chunk_len = c->chunk_len;
      sp.k_hi = N - 1;
    }
    HIPCHK(launch_src<T>(sp, nb, st));
  }
  HIPCHK(hipEventRecord(c->xev[0], st));
  HIPCHK(hipStreamWaitEvent(st2, c->xev[0], 0));
  HIPCHK(launch_upcont<T>(u, nb, st2, te[4], te[5]));
  HIPCHK(launch_surface<T>(a, nb, st2, te[6], te[7]));
  c->cur ^= 1;  // the new state is current
  advance_time(c);
  if (int rc = sample_rx<T>(c, st2, timed)) return rc;
  HIPCHK(hipEventRecord(c->xev[1], st2));
  if (timed) {
    // the march of a step is its lower and top launches (not the wait between them for the
    // previous step's stream2 work); the step is counted once
    if (split) c->spans.push_back({0, 0, te[0], te[1]});
    c->spans.push_back({0, 1, te[2], te[3]});
    c->spans.push_back({2, 1, te[4], te[5]});
    c->spans.push_back({1, 1, te[6], te[7]});
  }
  return TCSEM_OK;
}

// ---- graph replays ------------------------------------------------------------------------------
// Whether a receiver sample falls in the next k steps: the sampling bookkeeping, which depends on
// the time schedule only, run on a copy of the state.
bool events_ahead(tcsem_ctx *c, int k) {
  if (!c->have_rx) return false;
  const int n0 = c->n, st0 = c->steps, ne0 = c->ne, nb0 = c->nb;
  const double toT0 = c->toT, d10 = c->dt1, d20 = c->dt2;
  double tp0[2][3];
  std::memcpy(tp0, c->Tpof, sizeof tp0);
  std::vector<int> it;
  std::vector<double> l1;
  const bool has_b = c->rt.is_comp[3] || c->rt.is_comp[4] || c->rt.is_comp[5];
  bool fire = false;
  for (int q = 0; q < k && !fire; ++q) {
    advance_time(c);
    const double t0 = c->toT + c->dt1;
    for (int fld = 0; fld < 2 && !fire; ++fld) {
      if (fld == 0 && c->ne >= c->rt.Nt) continue;
      if (fld == 1 && (c->nb >= c->rt.Nt || !has_b)) continue;
      Event ev{};
      fire = sample_group(c, fld, fld == 0 ? t0 : t0 + 0.5 * c->dt2, it, l1, ev);
    }
  }
  c->n = n0;
  c->steps = st0;
  c->ne = ne0;
  c->nb = nb0;
  c->toT = toT0;
  c->dt1 = d10;
  c->dt2 = d20;
  std::memcpy(c->Tpof, tp0, sizeof tp0);
  return fire;
}

// Steps that a graph may replay: past the source (no amplitude, no source pass), no timing
// events, and no receiver sample among the next graph_k steps.
bool graph_ok(tcsem_ctx *c) {
  return c->graph_k > 0 && c->n >= 2 * (c->setup.n_src / 2) && !c->timing && !events_ahead(c, c->graph_k);
}

// The kernels of k steps as do_step issues them (steady state: no source pass, no gathers), with
// the step scalars read from d_gs; the first step's wait for the previous stream-2 work is left to
// the caller (issued before the capture).
template <typename T>
hipError_t enqueue_graph_steps(tcsem_ctx *c, int k) {
  const int nb = c->nbatch;
  hipStream_t st = c->stream, st2 = c->stream2;
  const UpcontArgs<T> &u = up_of<T>(c);
  hipError_t e;
  for (int q = 0; q < k; ++q) {
    MarchArgs<T> a = march_args<T>(c);
    a.gs = c->d_gs + 4 * q;
    if ((e = enqueue_march<T>(c, a, nb, st, q ? c->xev[1] : nullptr, nullptr)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->xev[0], st)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(st2, c->xev[0], 0)) != hipSuccess) return e;
    if ((e = launch_upcont<T>(u, nb, st2)) != hipSuccess) return e;
    if ((e = launch_surface<T>(a, nb, st2)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->xev[1], st2)) != hipSuccess) return e;
    c->cur ^= 1;
  }
  return hipStreamWaitEvent(st, c->xev[1], 0);  // join stream 2 into the capture
}

// graph_k steps as one graph replay (captured on first use for this starting state).
template <typename T>
int graph_steps(tcsem_ctx *c) {
  const int k = c->graph_k;
  hipStream_t st = c->stream;
  // the step scalars of the k steps, from a copy of the time state
  GraphScalars gs{};
  {
    const int n0 = c->n, st0 = c->steps;
    const double toT0 = c->toT, d10 = c->dt1, d20 = c->dt2;
    for (int q = 0; q < k; ++q) {
      MarchArgs<T> a{};
      step_scalars<T>(c, a);
      gs.v[4 * q] = a.lam;
      gs.v[4 * q + 1] = a.dt;
      gs.v[4 * q + 2] = a.dtb;
      gs.v[4 * q + 3] = a.amp;
      advance_time(c);
    }
    c->n = n0;
    c->steps = st0;
    c->toT = toT0;
    c->dt1 = d10;
    c->dt2 = d20;
  }
  if (c->nchunk > 1 && c->fuse) HIPCHK(ensure_ib(c, c->nchunk));  // no allocation inside a capture
  const tcsem_ctx::Graph *g = nullptr;
  for (const auto &x : c->graphs)
    if (x.cur == c->cur && x.ib_par == c->ib_par && x.ib_valid == (int)c->ib_valid && x.by == c->by && x.pf == c->pf &&
        x.nchunk == c->nchunk && x.chunk_len == c->chunk_len && x.fuse == c->fuse && x.s8 == (int)c->s8)
      g = &x;
  if (!g) {
    tcsem_ctx::Graph ng{c->cur, c->ib_par, (int)c->ib_valid, c->by,        c->pf, c->nchunk, c->chunk_len,
                        c->fuse, (int)c->s8, 0,          0,           0,     nullptr};
    const int ibn0 = c->ib_nchunk, ibl0 = c->ib_chunk_len;
    HIPCHK(hipStreamWaitEvent(st, c->xev[1], 0));
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const hipError_t e = enqueue_graph_steps<T>(c, k);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(st, &graph);
    ng.cur_after = c->cur;
    ng.ib_par_after = c->ib_par;
    ng.ib_valid_after = c->ib_valid;
    c->cur = ng.cur;
    c->ib_par = ng.ib_par;
    c->ib_valid = ng.ib_valid != 0;
    c->ib_nchunk = ibn0;
    c->ib_chunk_len = ibl0;
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      c->graph_k = 0;  // run the plain steps from here on
      return fail(TCSEM_EDEVICE, "step capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    const hipError_t e3 = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e3 != hipSuccess) {
      c->graph_k = 0;
      return fail(TCSEM_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(e3));
    }
    c->graphs.push_back(ng);
    g = &c->graphs.back();
  }
  HIPCHK(hipStreamWaitEvent(st, c->xev[1], 0));  // the previous step's stream-2 work
  HIPCHK(launch_put_scalars(c->d_gs, gs, k, st));
  HIPCHK(hipGraphLaunch(g->exec, st));
  HIPCHK(hipEventRecord(c->xev[1], st));  // the replay joined stream 2's work into stream
  c->cur = g->cur_after;
  c->ib_par = g->ib_par_after;
  c->ib_valid = g->ib_valid_after != 0;
  if (c->nchunk > 1) {
    c->ib_nchunk = c->nchunk;
    c->ib_chunk_len = c->chunk_len;
  }
  for (int q = 0; q < k; ++q) {  // host bookkeeping of the k steps (no sample fires)
    advance_time(c);
    if (int rc = sample_rx<T>(c, c->stream2, false)) return rc;
  }
  ++c->graph_replays;
  return TCSEM_OK;
}

// ---- two steps per lower pass ---------------------------------------------------------------------
// A pair of steps: k_march2 takes both steps' levels N-1 .. 3 (step A's level 2 too) in one pass
// over the old state; then, in stream order, step A's top launch (level 1, into the intermediate
// state), its upward continuation and surface level, and step B's top launch (levels 2 and 1,
// from the intermediate state); step B's upward continuation, surface level and receiver gathers
// go to stream2 and overlap the next pass.  The pass reads the old state at levels >= 1 only and
// writes the intermediate state at level 2 only, so it too runs beside the previous step's
// stream2 work.
// Whether the next two steps may run as a pair: pairs on, past the source (no amplitude, no source
// pass), and no receiver sample after the first of them (one after the second is taken from the
// new state as usual).
bool pair_ok(tcsem_ctx *c, long room) {
  return c->pair_by > 0 && room >= 2 && c->g.info.N >= 6 && c->n >= 2 * (c->setup.n_src / 2) &&
         !events_ahead(c, 1);
}

hipError_t ensure_mid(tcsem_ctx *c) {
  if (c->d_mid) return hipSuccess;
  const size_t mb = (size_t)5 * c->nslots * c->fstride * c->esz;
  hipError_t e = hipMalloc(&c->d_mid, mb);
  if (e != hipSuccess) {
    c->d_mid = nullptr;
    return e;
  }
  return launch_fill_zero(c->d_mid, mb, c->stream);  // cells outside the grid stay zero
}

// The kernels of a pair of steps on nb slots: a carries step A's scalars (old state -> new state
// pointers), sb step B's; te: null or 14 events (start/end of the pass, step A's top launch,
// chain and surface, step B's top launch, chain and surface).  Stream2's work of the pair is left
// unjoined.
template <typename T>
hipError_t enqueue_pair(tcsem_ctx *c, const MarchArgs<T> &a, const MarchArgs<T> &sb, int nb, const hipEvent_t *te) {
  static const hipEvent_t none[14] = {};
  if (!te) te = none;
  hipStream_t st = c->stream, st2 = c->stream2;
  const UpcontArgs<T> &u = up_of<T>(c);
  hipError_t e;
  if ((e = ensure_mid(c)) != hipSuccess) return e;
  const long long org = c->PX + c->xm, fs = c->fstride * c->nslots, ps = c->pstride * c->nslots;
  T *fm = (T *)c->d_mid, *pl = (T *)c->d_planes;
  PairArgs<T> p{};
  p.a = a;
  p.a.pf = c->pair_pf;
  p.mex = fm + org;
  p.mey = fm + fs + org;
  p.mez = fm + 2 * fs + org;
  p.mbx = fm + 3 * fs + org;
  p.mby = fm + 4 * fs + org;
  p.cs2 = pl + 5 * ps + org;
  p.cc2 = pl + 6 * ps + org;
  p.lamB = sb.lam;
  p.dtB = sb.dt;
  p.dtbB = sb.dtb;
  p.rows = c->pair_by;
  p.ntx = c->pair_ntx;
  p.nty = c->pair_nty;
  p.nper = c->pair_nper;
  // top launch and surface of step A: old state -> intermediate state
  MarchArgs<T> ta = a;
  ta.ex = p.mex;
  ta.ey = p.mey;
  ta.ez = p.mez;
  ta.bx = p.mbx;
  ta.by = p.mby;
  ta.k_hi = 1;
  ta.carry = 1;
  // top launch (levels 2, 1) and surface of step B: intermediate state -> new state
  MarchArgs<T> tb = a;
  tb.exi = p.mex;
  tb.eyi = p.mey;
  tb.ezi = p.mez;
  tb.bxi = p.mbx;
  tb.byi = p.mby;
  tb.lam = sb.lam;
  tb.dt = sb.dt;
  tb.dtb = sb.dtb;
  tb.amp = 0.0;
  tb.k_hi = 2;
  tb.carry = 1;
  tb.cs = p.cs2;
  tb.cc = p.cc2;
  if ((e = launch_march2<T>(p, nb, st, te[0], te[1])) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(st, c->xev[1], 0)) != hipSuccess) return e;  // the previous step's stream2 work
  if ((e = launch_march<T>(ta, nb, st, te[2], te[3])) != hipSuccess) return e;
  if ((e = launch_upcont<T>(u, nb, st, te[4], te[5], nullptr, true)) != hipSuccess) return e;
  if ((e = launch_surface<T>(ta, nb, st, te[6], te[7])) != hipSuccess) return e;
  if ((e = launch_march<T>(tb, nb, st, te[8], te[9])) != hipSuccess) return e;
  if ((e = hipEventRecord(c->xev[0], st)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(st2, c->xev[0], 0)) != hipSuccess) return e;
  if ((e = launch_upcont<T>(u, nb, st2, te[10], te[11])) != hipSuccess) return e;
  return launch_surface<T>(tb, nb, st2, te[12], te[13]);
}

template <typename T>
int do_pair(tcsem_ctx *c) {
  MarchArgs<T> a = march_args<T>(c);  // old state -> new state, step A's scalars
  step_scalars<T>(c, a);
  MarchArgs<T> sb{};  // step B's scalars: the time state after step A
  {
    const int n0 = c->n, st0 = c->steps;
    const double toT0 = c->toT, d10 = c->dt1, d20 = c->dt2;
    advance_time(c);
    step_scalars<T>(c, sb);
    c->n = n0;
    c->steps = st0;
    c->toT = toT0;
    c->dt1 = d10;
    c->dt2 = d20;
  }
  // a pair is timed when one of its steps would be (every kTimingStride-th step)
  const bool timed = c->timing && (c->steps % kTimingStride == 0 || (c->steps + 1) % kTimingStride == 0);
  hipEvent_t te[14] = {};
  if (timed) {
    if (c->spans.size() > 50000 && (rc_collect(c))) return TCSEM_EDEVICE;
    for (auto &e : te)
      if (!(e = timing_event(c))) return fail(TCSEM_EDEVICE, "timing event creation failed");
  }
  HIPCHK(enqueue_pair<T>(c, a, sb, c->nbatch, timed ? te : nullptr));
  c->cur ^= 1;
  c->ib_valid = false;
  advance_time(c);
  if (int rc = sample_rx<T>(c, c->stream2, false)) return rc;  // no sample fires after step A
  advance_time(c);
  if (int rc = sample_rx<T>(c, c->stream2, timed)) return rc;
  HIPCHK(hipEventRecord(c->xev[1], c->stream2));
  ++c->pairs_done;
  if (timed) {
    // the march of the two steps: the pass and both top launches (two steps counted)
    c->spans.push_back({0, 0, te[0], te[1]});
    c->spans.push_back({0, 1, te[2], te[3]});
    c->spans.push_back({0, 1, te[8], te[9]});
    c->spans.push_back({2, 1, te[4], te[5]});
    c->spans.push_back({2, 1, te[10], te[11]});
    c->spans.push_back({1, 1, te[6], te[7]});
    c->spans.push_back({1, 1, te[12], te[13]});
    c->spans.push_back({4, 1, te[0], te[1]});
  }
  return TCSEM_OK;
}

// One step, or graph_k steps as a graph replay when allowed and at most `room` steps may be taken.
// Returns the steps taken (negative: error code).
template <typename T>
int advance(tcsem_ctx *c, long room) {
  if (room >= c->graph_k && graph_ok(c)) {
    const int rc = graph_steps<T>(c);
    return rc ? rc : c->graph_k;
  }
  if (pair_ok(c, room)) {
    const int rc = do_pair<T>(c);
    return rc ? rc : 2;
  }
  const int rc = do_step<T>(c);
  return rc ? rc : 1;
}

// Orders everything issued on stream2 before what is issued next on stream; every entry point
// that returns after stepping calls it, so host-ordered operations see a whole step.
int join_streams(tcsem_ctx *c) {
  HIPCHK(hipEventRecord(c->xev[1], c->stream2));
  HIPCHK(hipStreamWaitEvent(c->stream, c->xev[1], 0));
  return TCSEM_OK;
}

bool sampling_done(const tcsem_ctx *c) {  // main.cu:219
  if (!c->have_rx) return false;
  return !(c->ne < c->rt.Nt || (c->nb < c->rt.Nt && c->nefcmp < c->ncomp));
}

int check_ctx(const tcsem_ctx *c) {
  if (!c) return fail(TCSEM_EINVAL, "null context");
  return TCSEM_OK;
}

}  // namespace

// ==================================================================================================
// C ABI
// ==================================================================================================
// Times nbatch-slot steps as they run in production -- the volume march on the main stream
// beside the upward continuation and surface level of the previous step on the second stream,
// on zeroed fields without a source -- for each candidate march tiling, and returns the fastest
// (its time per step).  Timing the march alone would favour tilings that lose more bandwidth
// to the concurrent air-plane chain.
template <typename T>
int autotune_march(tcsem_ctx *c, int nbatch, int nsteps, int *best_by, int *best_pf, int *best_nch, int *best_fuse,
                   int *best_s8, double *best_ms) {
  const int N = c->g.info.N, nlev = N - 2;
  std::vector<int> rows_c = c->prec == 64 ? std::vector<int>{4, 8, 10, 12} : std::vector<int>{4, 8, 10, 16};
  std::vector<int> nch_c = {1};
  if (N >= 6)
    for (int z : {2, 3, 4, 6, 8})
      if (z <= std::max(1, nlev / 8)) nch_c.push_back(z);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(TCSEM_EDEVICE, "event creation failed");
  }
  hipStream_t st = c->stream, st2 = c->stream2;
  const UpcontArgs<T> &u = up_of<T>(c);
  const int cur0 = c->cur;
  auto one_step = [&]() -> hipError_t {
    MarchArgs<T> a = march_args<T>(c);
    a.lam = 1.0;
    a.dt = 1e-6;
    a.dtb = 1e-6;
    a.amp = 0.0;
    hipError_t e = enqueue_march<T>(c, a, nbatch, st, c->xev[1], nullptr);
    if (e == hipSuccess) e = hipEventRecord(c->xev[0], st);
    if (e == hipSuccess) e = hipStreamWaitEvent(st2, c->xev[0], 0);
    if (e == hipSuccess) e = launch_upcont<T>(u, nbatch, st2, nullptr, nullptr);
    if (e == hipSuccess) e = launch_surface<T>(a, nbatch, st2, nullptr, nullptr);
    if (e == hipSuccess) e = hipEventRecord(c->xev[1], st2);
    c->cur ^= 1;
    return e;
  };
  struct Trial {
    double ms;
    int by, pf, nch, fuse, s8;
  };
  // time nsteps steps of one tiling (after an untimed step: the first launches of the tiling)
  auto time_tiling = [&](int by, int pf, int nch, int fuse, int s8, double *ms_out) -> hipError_t {
    apply_tiling(c, by, pf, nch);
    c->fuse = fuse;
    c->s8 = s8 != 0;
    hipError_t e = one_step();
    if (e == hipSuccess) e = hipStreamWaitEvent(st, c->xev[1], 0);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int q = 0; q < nsteps && e == hipSuccess; ++q) e = one_step();
    if (e == hipSuccess) e = hipStreamWaitEvent(st, c->xev[1], 0);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    *ms_out = (double)ms / nsteps;
    return e;
  };
  std::vector<Trial> trials;
  int rc = TCSEM_OK;
  const int fuse0 = c->fuse;
  const bool s80 = c->s8;
  for (int by : rows_c)
    for (int pf = 1; pf <= 4 && rc == TCSEM_OK; ++pf)
      for (int nch : nch_c)
        for (int fuse = 1; fuse >= 0 && rc == TCSEM_OK; --fuse)
          for (int s8 = c->d_sig8 ? 1 : 0; s8 >= 0 && rc == TCSEM_OK; --s8) {  // the index, if any, and not
            apply_tiling(c, by, pf, nch);
            if (c->nchunk == 0 && nch > 1) continue;    // clamped to one pass: already timed
            if (c->nchunk == 0 && fuse == 0) continue;  // no chunks: nothing to sum
            double ms = 0.0;
            const hipError_t e = time_tiling(by, pf, nch, fuse, s8, &ms);
            if (e != hipSuccess) {
              rc = fail(TCSEM_EDEVICE, "autotune: %s", hipGetErrorString(e));
              break;
            }
            trials.push_back({ms, c->by, pf, c->nchunk > 1 ? c->nchunk : 1, fuse, s8});
          }
  // the four fastest are timed twice more, interleaved, and ranked by the mean of their three
  // timings: one pass of short trials lets noise pick a tiling a few percent slower in production
  if (rc == TCSEM_OK && trials.size() > 1) {
    std::sort(trials.begin(), trials.end(), [](const Trial &x, const Trial &y) { return x.ms < y.ms; });
    if (trials.size() > 4) trials.resize(4);
    std::vector<double> sum(trials.size());
    for (size_t t = 0; t < trials.size(); ++t) sum[t] = trials[t].ms;
    for (int pass = 0; pass < 2 && rc == TCSEM_OK; ++pass)
      for (size_t t = 0; t < trials.size() && rc == TCSEM_OK; ++t) {
        double ms = 0.0;
        const hipError_t e = time_tiling(trials[t].by, trials[t].pf, trials[t].nch, trials[t].fuse, trials[t].s8, &ms);
        if (e != hipSuccess) rc = fail(TCSEM_EDEVICE, "autotune: %s", hipGetErrorString(e));
        sum[t] += ms;
      }
    for (size_t t = 0; t < trials.size(); ++t) trials[t].ms = sum[t] / 3.0;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->cur = cur0;
  c->fuse = fuse0;
  c->s8 = s80;
  c->ib_valid = false;
  if (rc) return rc;
  // the fastest, except that among tilings within 1% of it (timing noise) the one with the fewest
  // z-chunks wins: chunks add the partial sums of Bz (a k_zsum pass over Bx and By, 2 more words
  // per cell, or the chunk sums of B accumulated in FP64 by the march)
  double best = 1e300;
  for (const Trial &t : trials) best = std::min(best, t.ms);
  const Trial *pick = nullptr;
  for (const Trial &t : trials)
    if (t.ms <= 1.01 * best && (!pick || t.nch < pick->nch || (t.nch == pick->nch && t.ms < pick->ms))) pick = &t;
  if (!pick) return fail(TCSEM_EDEVICE, "autotune: no tiling timed");
  *best_by = pick->by;
  *best_pf = pick->pf;
  *best_nch = pick->nch;
  *best_fuse = pick->fuse;
  *best_s8 = pick->s8;
  *best_ms = pick->ms;
  return TCSEM_OK;
}

// Times pairs of steps as they run in production (the two-step pass, then the top levels of both
// steps, step B's chain and surface beside the next pass) for each pass shape, with the one-step
// tiling already applied for the top launches; returns the fastest shape and its time per step
// (rows 0: none timed).  The shapes are timed in two interleaved rounds and each keeps its faster
// time, so that one noisy sample does not decide between shapes a few percent apart.
template <typename T>
int autotune_pairs(tcsem_ctx *c, int nbatch, int nsteps, int *best_by, int *best_pf, double *best_ms) {
  *best_by = 0;
  *best_pf = 1;
  *best_ms = 1e300;
  if (c->g.info.N < 6) return TCSEM_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(TCSEM_EDEVICE, "event creation failed");
  }
  hipStream_t st = c->stream;
  const int cur0 = c->cur, pby0 = c->pair_by, ppf0 = c->pair_pf;
  auto one_pair = [&]() -> hipError_t {
    MarchArgs<T> a = march_args<T>(c);
    a.lam = 1.0;
    a.dt = 1e-6;
    a.dtb = 1e-6;
    a.amp = 0.0;
    hipError_t e = enqueue_pair<T>(c, a, a, nbatch, nullptr);
    if (e == hipSuccess) e = hipEventRecord(c->xev[1], c->stream2);
    c->cur ^= 1;
    return e;
  };
  const int npair = std::max(1, nsteps / 2);
  int rc = TCSEM_OK;
  struct Shape {
    int by, pf;
    double ms;
  };
  std::vector<Shape> shapes;
  for (int by : {12, 14, 16, 11, 10, 8})
    for (int pf = 1; pf <= 2; ++pf)
      if (pair_rows_supported(c->prec, by)) shapes.push_back({by, pf, 1e300});
  for (int round = 0; round < 2 && rc == TCSEM_OK; ++round)
    for (Shape &sh : shapes) {
      const int by = sh.by, pf = sh.pf;
      pair_tiling(c, by, pf);
      hipError_t e = one_pair();  // the first launches of the shape, untimed
      if (e == hipSuccess) e = hipStreamWaitEvent(st, c->xev[1], 0);
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int q = 0; q < npair && e == hipSuccess; ++q) e = one_pair();
      if (e == hipSuccess) e = hipStreamWaitEvent(st, c->xev[1], 0);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(TCSEM_EDEVICE, "autotune: %s", hipGetErrorString(e));
        break;
      }
      sh.ms = std::min(sh.ms, (double)ms / (2.0 * npair));
    }
  for (const Shape &sh : shapes)
    if (rc == TCSEM_OK && sh.ms < *best_ms) {
      *best_ms = sh.ms;
      *best_by = sh.by;
      *best_pf = sh.pf;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->cur = cur0;
  pair_tiling(c, pby0, ppf0);
  return rc;
}

// ---- model batch ----------------------------------------------------------------------------------
// Models of uneven grids, one transmitter each (every member a context begun with one slot), stepped
// together: each step's volume march, air-plane chain and surface level run as one launch over all
// members, from per-member argument records (2 parities of MarchArgs: the member's own fields,
// tables, sizes and tiling; one UpcontArgs each), with the step scalars of every member in a
// device table (field-independent, precomputed per call).  The source pass is one launch over the
// members too; receiver gathers stay per member.
struct tcsem_batch {
  std::vector<tcsem_ctx *> m;
  int prec = 32;
  int by = 0, pf = 0;  // autotuned march tiling (0: heuristic)
  bool s8 = true;      // the march reads the members' conductivity indices (when all have one)
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t xev[2] = {}, done = nullptr;
  char *d_geo = nullptr;     // 2 x nmodel MarchArgs (parity-major)
  char *d_ugeo = nullptr;    // nmodel UpcontArgs
  double *d_scal = nullptr;  // steps x nmodel x (lam, dt, dtb, amp)
  int scal_cap = 0;
  int2 *d_work = nullptr;    // march work list: (member, tile) per workgroup
  size_t work_cap = 0;
  int work_len = 0;          // workgroups of a march launch
  std::vector<char> active;  // per member: still sampling
  int steps = 0;
  long nenq = 0;             // batch steps enqueued (live or not)
  // timing, as for a context: kernel-recorded events every kTimingStride-th step, folded on read
  bool timing = false;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  std::vector<tcsem_ctx::Span> spans;  // kinds 0 march, 1 surface, 2 air-plane chain, 3 two-step pass
  double tacc[4] = {0, 0, 0, 0};
  long tn[4] = {0, 0, 0, 0};
  // Two steps per pass (prefetch 1): tcsem_batch_set_pairs' rows (-1: none), the autotuner's (-1:
  // not tuned), the rows of the current records (0: off); per-member pass records (2 parities)
  // and the top launches' records (step A's then step B's, 2 parities each); the pass work list.
  int pair_set = -1, pair_tuned = -1, pair_by = 0;
  char *d_pgeo = nullptr;  // 2 x nmodel PairArgs
  char *d_tgeo = nullptr;  // 4 x nmodel MarchArgs
  int2 *d_pwork = nullptr;
  size_t pwork_cap = 0;
  int pwork_len = 0;
  long pairs_done = 0;
};

namespace {

int batch_fail_if(bool bad, const char *msg) { return bad ? fail(TCSEM_ESTATE, "%s", msg) : TCSEM_OK; }

// Common tiling of a batch step: the autotuned one, else 10-row tiles (FP32) or 8-row tiles
// (FP64) at prefetch 1; TCSEM_MARCH_ROWS / _PF override either.
void batch_tiling(const tcsem_batch *b, int &by, int &pf) {
  if (b->by) {
    by = b->by;
    pf = b->pf;
  } else {  // the batch autotuner's pick on the 64-model workload
    by = b->prec == 64 ? 8 : 10;
    pf = 1;
  }
  if (const char *e = std::getenv("TCSEM_MARCH_ROWS")) {
    const int v = std::atoi(e);
    if (v == 4 || v == 8 || v == 10 || v == 12 || v == 16) by = v;
    if (supported_rows(b->prec, by) != by)
      std::fprintf(stderr, "tcsem: TCSEM_MARCH_ROWS=%d has no FP%d march; running %d rows\n", by, b->prec,
                   supported_rows(b->prec, by));
  }
  if (const char *e = std::getenv("TCSEM_MARCH_PF")) {
    const int v = std::atoi(e);
    if (v >= 1 && v <= 4) pf = v;
  }
}

// Builds and uploads the member records for tiling (by, pf) and returns the launch records: the
// largest sizes over the members (grid and LDS extents), the common tiling and level range.
// The batch's streams must be idle (the records are overwritten).
// A work list of (member, tile) per workgroup for tiles[q] tiles of each active member: workgroup
// w runs on XCD w mod 8, so list x of the 8 per-XCD lists fills workgroups x, x + 8, ...  Each
// active member's tiles go to one XCD where possible, which keeps the halo rows shared by its
// neighbouring tiles in that XCD's L2; a member with more tiles than half an XCD's share is cut
// into row-major bands of that size.  Pieces go, largest first, to the XCD with the least work
// (tiles x levels).  Returns the list's length (a multiple of 8), uploaded to *d (grown as needed).
int batch_work_list(tcsem_batch *b, const std::vector<int> &tiles, int2 **d, size_t *cap, int *len) {
  const int nm = (int)b->m.size();
  long long total = 0;
  for (int q = 0; q < nm; ++q)
    if (b->active[q]) total += tiles[q];
  const long long chunk = std::max(1LL, (total + 15) / 16);
  struct Piece {
    int member, t0, n;
    long long work;
  };
  std::vector<Piece> pieces;
  for (int q = 0; q < nm; ++q) {
    if (!b->active[q]) continue;
    const int nt = tiles[q];
    for (int t0 = 0; t0 < nt; t0 += (int)chunk) {
      const int n = (int)std::min<long long>(chunk, nt - t0);
      pieces.push_back({q, t0, n, (long long)n * b->m[q]->g.info.N});
    }
  }
  std::stable_sort(pieces.begin(), pieces.end(), [](const Piece &x, const Piece &y) { return x.work > y.work; });
  std::vector<std::vector<int2>> lists(8);
  long long load[8] = {0};
  for (const Piece &pc : pieces) {
    const int x = (int)(std::min_element(load, load + 8) - load);
    for (int t = pc.t0; t < pc.t0 + pc.n; ++t) lists[x].push_back(make_int2(pc.member, t));
    load[x] += pc.work;
  }
  size_t n = 1;
  for (const auto &l : lists) n = std::max(n, l.size());
  std::vector<int2> tbl(8 * n, make_int2(-1, -1));
  for (int x = 0; x < 8; ++x)
    for (size_t i = 0; i < lists[x].size(); ++i) tbl[8 * i + x] = lists[x][i];
  if (tbl.size() > *cap) {
    (void)hipFree(*d);
    *d = nullptr;
    *cap = 0;
    if (hipMalloc((void **)d, tbl.size() * sizeof(int2)) != hipSuccess) return fail(TCSEM_ENOMEM, "work list allocation");
    *cap = tbl.size();
  }
  HIPCHK(hipMemcpy(*d, tbl.data(), tbl.size() * sizeof(int2), hipMemcpyHostToDevice));
  *len = (int)tbl.size();
  return TCSEM_OK;
}

int pass_tiles_x(int L) { return (L + kMarchLanes - 5) / (kMarchLanes - 4); }
int pass_tiles_y(int M, int by) { return (M + by - 5) / (by - 4); }

// pby > 0: also the records of two steps per pass with pby-row tiles (each member's intermediate
// state allocated on first use).
template <typename T>
int batch_records(tcsem_batch *b, int by, int pf, MarchArgs<T> &la, UpcontArgs<T> &ul, int pby = 0) {
  const int nm = (int)b->m.size();
  std::vector<MarchArgs<T>> rec(2 * (size_t)nm);
  std::vector<UpcontArgs<T>> urec(nm);
  la = MarchArgs<T>{};
  ul = UpcontArgs<T>{};
  for (int q = 0; q < nm; ++q) {
    tcsem_ctx *c = b->m[q];
    apply_tiling(c, by, pf, 1);
    by = c->by;  // the row counts of the member's precision (apply_tiling)
    const int cur0 = c->cur;
    for (int p = 0; p < 2; ++p) {
      c->cur = cur0 ^ p;
      MarchArgs<T> a = march_args<T>(c);
      if (!b->active[q]) a.nper = a.sper = 0;  // every workgroup of the member returns at once
      if (b->active[q] && c->has_fix) {  // the member's source pass (its amplitude from the step table)
        a.fixe = c->d_fixe;
        a.fixb = c->d_fixb;
        a.fixg = c->d_fixg;
        a.fixidx = c->d_fixidx;
        a.fixn = 1;
      }
      rec[(size_t)p * nm + q] = a;
    }
    c->cur = cur0;
    UpcontArgs<T> u = up_of<T>(c);
    if (!b->active[q]) u.L = u.M = u.M1 = u.H = u.nyr = 0;
    urec[q] = u;
    const tcsem_grid_info &I = c->g.info;
    la.L = std::max(la.L, I.L);
    la.M = std::max(la.M, I.M);
    la.N = std::max(la.N, I.N);
    la.sper = std::max(la.sper, c->sper);
    ul.L = std::max(ul.L, u.L);
    ul.M = std::max(ul.M, u.M);
    ul.M1 = std::max(ul.M1, u.M1);
    ul.H = std::max(ul.H, u.H);
    ul.nyr = std::max(ul.nyr, u.nyr);
    ul.px.m = std::max(ul.px.m, u.px.m);
    ul.py.m = std::max(ul.py.m, u.py.m);
  }
  la.rows = by;
  la.pf = pf;
  la.nchunk = 0;
  // the march reads the 8-bit conductivity index when every member's record carries one (the
  // launch record's pointer is then only that flag)
  la.sig8 = nullptr;
  if (b->s8 && std::all_of(b->m.begin(), b->m.end(), [](const tcsem_ctx *c) { return c->d_sig8 != nullptr; }))
    la.sig8 = b->m[0]->d_sig8;
  for (int q = 0; q < nm; ++q) {  // every member record as the launch reads it (a member's own choice aside)
    const tcsem_ctx *c = b->m[q];
    for (int p = 0; p < 2; ++p) rec[(size_t)p * nm + q].sig8 = la.sig8 ? c->d_sig8 + c->PX + c->xm : nullptr;
  }
  la.k_hi = la.N - 1;
  la.k_lo = 1;
  ul.tiled = 1;
  // the march's work list
  {
    std::vector<int> tiles(nm);
    for (int q = 0; q < nm; ++q) tiles[q] = b->m[q]->ntx * b->m[q]->nty;
    if (int rc = batch_work_list(b, tiles, &b->d_work, &b->work_cap, &b->work_len)) return rc;
    la.work = b->d_work;
    la.nper = b->work_len / 8;  // the march grid: 8 x len workgroups, one row
  }
  b->pair_by = 0;
  if (pby > 0) {  // two steps per pass: the pass and top-launch records, the pass work list
    std::vector<PairArgs<T>> prec(2 * (size_t)nm);
    std::vector<MarchArgs<T>> trec(4 * (size_t)nm);
    std::vector<int> tiles(nm);
    for (int q = 0; q < nm; ++q) {
      tcsem_ctx *c = b->m[q];
      HIPCHK(ensure_mid(c));
      HIPCHK(hipStreamSynchronize(c->stream));
      const long long org = c->PX + c->xm, fs = c->fstride * c->nslots, ps = c->pstride * c->nslots;
      T *fm = (T *)c->d_mid, *pl = (T *)c->d_planes;
      const tcsem_grid_info &I = c->g.info;
      tiles[q] = pass_tiles_x(I.L) * pass_tiles_y(I.M, pby);
      for (int p = 0; p < 2; ++p) {
        const MarchArgs<T> &r = rec[(size_t)p * nm + q];
        PairArgs<T> pr{};
        pr.a = r;
        pr.mex = fm + org;
        pr.mey = fm + fs + org;
        pr.mez = fm + 2 * fs + org;
        pr.mbx = fm + 3 * fs + org;
        pr.mby = fm + 4 * fs + org;
        pr.cs2 = pl + 5 * ps + org;
        pr.cc2 = pl + 6 * ps + org;
        pr.rows = pby;
        pr.ntx = pass_tiles_x(I.L);
        pr.nty = pass_tiles_y(I.M, pby);
        prec[(size_t)p * nm + q] = pr;
        MarchArgs<T> ta = r;  // step A's top launch and surface: old state -> intermediate state
        ta.ex = pr.mex;
        ta.ey = pr.mey;
        ta.ez = pr.mez;
        ta.bx = pr.mbx;
        ta.by = pr.mby;
        trec[(size_t)p * nm + q] = ta;
        MarchArgs<T> tb = r;  // step B's top launch and surface: intermediate state -> new state
        tb.exi = pr.mex;
        tb.eyi = pr.mey;
        tb.ezi = pr.mez;
        tb.bxi = pr.mbx;
        tb.byi = pr.mby;
        tb.cs = pr.cs2;
        tb.cc = pr.cc2;
        trec[(size_t)(2 + p) * nm + q] = tb;
      }
    }
    if (int rc = batch_work_list(b, tiles, &b->d_pwork, &b->pwork_cap, &b->pwork_len)) return rc;
    HIPCHK(hipMemcpy(b->d_pgeo, prec.data(), prec.size() * sizeof(PairArgs<T>), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_tgeo, trec.data(), trec.size() * sizeof(MarchArgs<T>), hipMemcpyHostToDevice));
    b->pair_by = pby;
  }
  HIPCHK(hipMemcpy(b->d_geo, rec.data(), rec.size() * sizeof(MarchArgs<T>), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_ugeo, urec.data(), urec.size() * sizeof(UpcontArgs<T>), hipMemcpyHostToDevice));
  return TCSEM_OK;
}

int batch_scal_reserve(tcsem_batch *b, int nsteps) {
  if (nsteps <= b->scal_cap) return TCSEM_OK;
  (void)hipFree(b->d_scal);
  b->d_scal = nullptr;
  b->scal_cap = 0;
  const size_t bytes = (size_t)nsteps * b->m.size() * 4 * sizeof(double);
  if (hipMalloc((void **)&b->d_scal, bytes) != hipSuccess) return fail(TCSEM_ENOMEM, "step-scalar table allocation");
  b->scal_cap = nsteps;
  return TCSEM_OK;
}

int batch_collect(tcsem_batch *b);

// Waits for the members' own streams and the batch's, so records and tables can be rewritten.
int batch_quiesce(tcsem_batch *b) {
  for (tcsem_ctx *c : b->m) {
    HIPCHK(hipStreamSynchronize(c->stream2));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream2));
  HIPCHK(hipStreamSynchronize(b->stream));
  return batch_collect(b);
}

// One batch step on the batch's streams, as do_step for one context: lower march of every member,
// (after the previous step's stream2 work) the top march, the members' source pass; on stream2
// the air-plane chain, the surface level and the members' receiver gathers.  par: record parity.
template <typename T>
int batch_enqueue(tcsem_batch *b, const MarchArgs<T> &la, const UpcontArgs<T> &ul, int par, int row,
                  bool live) {
  const int nm = (int)b->m.size();
  const MarchArgs<T> *geo = (const MarchArgs<T> *)b->d_geo + (size_t)par * nm;
  const UpcontArgs<T> *ugeo = (const UpcontArgs<T> *)b->d_ugeo;
  hipStream_t st = b->stream, st2 = b->stream2;
  MarchArgs<T> lo = la, hi = la;
  lo.gscal = hi.gscal = b->d_scal + (size_t)row * nm * 4;
  lo.k_lo = 2;
  lo.carry = 2;
  hi.k_hi = 1;
  hi.k_lo = 1;
  hi.carry = 1;
  const bool timed = b->timing && b->nenq % kTimingStride == 0;
  hipEvent_t te[8] = {};
  if (timed)
    for (auto &e : te) {
      if (b->evused == b->evpool.size()) {
        hipEvent_t x = nullptr;
        if (hipEventCreate(&x) != hipSuccess) return fail(TCSEM_EDEVICE, "timing event creation failed");
        b->evpool.push_back(x);
      }
      e = b->evpool[b->evused++];
    }
  ++b->nenq;
  HIPCHK(launch_march<T>(lo, 1, st, te[0], te[1], geo));  // one grid row: the work list
  HIPCHK(hipStreamWaitEvent(st, b->xev[1], 0));
  HIPCHK(launch_march<T>(hi, 1, st, te[2], te[3], geo));
  if (live) {  // one source pass for every member with a source cell this step
    bool src = false;
    for (int q = 0; q < nm && !src; ++q) {
      tcsem_ctx *c = b->m[q];
      if (!b->active[q] || !c->has_fix) continue;
      MarchArgs<T> a{};
      step_scalars<T>(c, a);
      src = a.amp != 0.0;
    }
    if (src) HIPCHK(launch_src<T>(hi, nm, st, geo));
  }
  HIPCHK(hipEventRecord(b->xev[0], st));
  HIPCHK(hipStreamWaitEvent(st2, b->xev[0], 0));
  HIPCHK(launch_upcont<T>(ul, nm, st2, te[4], te[5], ugeo));
  HIPCHK(launch_surface<T>(hi, nm, st2, te[6], te[7], geo));
  if (timed) {
    b->spans.push_back({0, 0, te[0], te[1]});
    b->spans.push_back({0, 1, te[2], te[3]});
    b->spans.push_back({2, 1, te[4], te[5]});
    b->spans.push_back({1, 1, te[6], te[7]});
  }
  if (live)
    for (int q = 0; q < nm; ++q) {
      tcsem_ctx *c = b->m[q];
      c->cur ^= 1;
      advance_time(c);
      if (b->active[q])
        if (int rc = sample_rx<T>(c, st2, false)) return rc;
    }
  HIPCHK(hipEventRecord(b->xev[1], st2));
  return TCSEM_OK;
}

// Two batch steps as one two-step pass (rows k and k+1 of the scalar table, record parity par),
// then each step's top launch, chain and surface over every member, as do_pair issues them for a
// context.  live: advance the members' schedules and sample (no sample fires after the first step).
template <typename T>
int batch_enqueue_pair(tcsem_batch *b, const MarchArgs<T> &la, const UpcontArgs<T> &ul, int par, int row,
                       bool live) {
  const int nm = (int)b->m.size();
  const MarchArgs<T> *tga = (const MarchArgs<T> *)b->d_tgeo + (size_t)par * nm;
  const MarchArgs<T> *tgb = (const MarchArgs<T> *)b->d_tgeo + (size_t)(2 + par) * nm;
  const PairArgs<T> *pg = (const PairArgs<T> *)b->d_pgeo + (size_t)par * nm;
  const UpcontArgs<T> *ugeo = (const UpcontArgs<T> *)b->d_ugeo;
  hipStream_t st = b->stream, st2 = b->stream2;
  PairArgs<T> pl{};
  pl.a = la;
  pl.a.work = b->d_pwork;
  pl.a.pf = 1;
  pl.a.gscal = b->d_scal + (size_t)row * nm * 4;
  pl.gscalB = b->d_scal + (size_t)(row + 1) * nm * 4;
  pl.rows = b->pair_by;
  pl.nper = b->pwork_len / 8;
  MarchArgs<T> ta = la, tb = la;
  ta.gscal = pl.a.gscal;
  tb.gscal = pl.gscalB;
  ta.k_hi = 1;
  ta.k_lo = 1;
  ta.carry = 1;
  tb.k_hi = 2;
  tb.k_lo = 1;
  tb.carry = 1;
  const bool timed = b->timing && (b->nenq % kTimingStride == 0 || (b->nenq + 1) % kTimingStride == 0);
  hipEvent_t te[14] = {};
  if (timed)
    for (auto &e : te) {
      if (b->evused == b->evpool.size()) {
        hipEvent_t x = nullptr;
        if (hipEventCreate(&x) != hipSuccess) return fail(TCSEM_EDEVICE, "timing event creation failed");
        b->evpool.push_back(x);
      }
      e = b->evpool[b->evused++];
    }
  b->nenq += 2;
  HIPCHK(launch_march2<T>(pl, 1, st, te[0], te[1], pg));
  HIPCHK(hipStreamWaitEvent(st, b->xev[1], 0));
  HIPCHK(launch_march<T>(ta, 1, st, te[2], te[3], tga));
  HIPCHK(launch_upcont<T>(ul, nm, st, te[4], te[5], ugeo, true));
  HIPCHK(launch_surface<T>(ta, nm, st, te[6], te[7], tga));
  HIPCHK(launch_march<T>(tb, 1, st, te[8], te[9], tgb));
  HIPCHK(hipEventRecord(b->xev[0], st));
  HIPCHK(hipStreamWaitEvent(st2, b->xev[0], 0));
  HIPCHK(launch_upcont<T>(ul, nm, st2, te[10], te[11], ugeo));
  HIPCHK(launch_surface<T>(tb, nm, st2, te[12], te[13], tgb));
  if (timed) {
    b->spans.push_back({0, 0, te[0], te[1]});
    b->spans.push_back({0, 1, te[2], te[3]});
    b->spans.push_back({0, 1, te[8], te[9]});
    b->spans.push_back({2, 1, te[4], te[5]});
    b->spans.push_back({2, 1, te[10], te[11]});
    b->spans.push_back({1, 1, te[6], te[7]});
    b->spans.push_back({1, 1, te[12], te[13]});
    b->spans.push_back({3, 1, te[0], te[1]});
  }
  if (live) {
    for (int q = 0; q < nm; ++q) {
      tcsem_ctx *c = b->m[q];
      c->cur ^= 1;
      for (int h = 0; h < 2; ++h) {
        advance_time(c);
        if (b->active[q])
          if (int rc = sample_rx<T>(c, st2, false)) return rc;
      }
    }
    ++b->pairs_done;
  }
  HIPCHK(hipEventRecord(b->xev[1], st2));
  return TCSEM_OK;
}

// Whether the batch's next two steps may run as a pair: every member still sampling is past its
// source wavelet and takes no receiver sample after the first of the two steps.
bool batch_pair_ok(tcsem_batch *b) {
  for (size_t q = 0; q < b->m.size(); ++q) {
    if (!b->active[q]) continue;
    tcsem_ctx *c = b->m[q];
    if (c->n < 2 * (c->setup.n_src / 2) || events_ahead(c, 1)) return false;
  }
  return true;
}

// Rows of the batch's two-step passes: tcsem_batch_set_pairs' choice, else TCSEM_MARCH_PAIRS
// ("0" or "rows[,1]"), else the autotuner's, else 12-row (FP64: 8-row) passes when the members'
// pass tiles give at least two workgroups per CU (0: one step at a time); measured on the 64
// uneven models, FP64 batches ran 6% faster with 8 rows than with 12.
int batch_pair_rows(tcsem_batch *b) {
  if (b->pair_set >= 0) return b->pair_set;
  if (const char *e = std::getenv("TCSEM_MARCH_PAIRS")) {
    int r = 0, q = 1;
    if (std::sscanf(e, "%d,%d", &r, &q) >= 1 && (r == 0 || (pair_rows_supported(b->prec, r) && q == 1))) return r;
  }
  if (b->pair_tuned >= 0) return b->pair_tuned;
  const int rows = b->prec == 64 ? 8 : 12;
  long long tiles = 0;
  for (size_t q = 0; q < b->m.size(); ++q)
    if (b->active[q])
      tiles += (long long)pass_tiles_x(b->m[q]->g.info.L) * pass_tiles_y(b->m[q]->g.info.M, rows);
  return tiles >= 2LL * device_cus() ? rows : 0;
}

// Orders the batch's work before what is issued next on the batch's and the members' streams.
int batch_join(tcsem_batch *b) {
  HIPCHK(hipEventRecord(b->xev[1], b->stream2));
  HIPCHK(hipStreamWaitEvent(b->stream, b->xev[1], 0));
  HIPCHK(hipEventRecord(b->done, b->stream));
  for (tcsem_ctx *c : b->m) {
    HIPCHK(hipStreamWaitEvent(c->stream, b->done, 0));
    HIPCHK(hipStreamWaitEvent(c->stream2, b->done, 0));
  }
  return TCSEM_OK;
}

// Waits for the recorded intervals of the batch and adds them to the per-kind totals.
int batch_collect(tcsem_batch *b) {
  if (b->spans.empty()) return TCSEM_OK;
  HIPCHK(hipStreamSynchronize(b->stream2));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (const auto &sp : b->spans) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, sp.a, sp.b));
    b->tacc[sp.kind] += ms;
    b->tn[sp.kind] += sp.count;
  }
  b->spans.clear();
  b->evused = 0;
  return TCSEM_OK;
}

template <typename T>
int batch_step_t(tcsem_batch *b, int nsteps) {
  const int nm = (int)b->m.size();
  if (int rc = batch_quiesce(b)) return rc;
  bool any = false;
  for (int q = 0; q < nm; ++q) {
    b->active[q] = !sampling_done(b->m[q]);
    any = any || b->active[q];
    b->m[q]->ib_valid = false;  // batch steps write no chunk sums
  }
  if (!any) return TCSEM_OK;  // every member has taken all its samples
  // every member's step scalars for the nsteps steps (the schedule is field-independent)
  if (int rc = batch_scal_reserve(b, nsteps)) return rc;
  std::vector<double> scal((size_t)nsteps * nm * 4, 0.0);
  for (int q = 0; q < nm; ++q) {
    tcsem_ctx *c = b->m[q];
    const int n0 = c->n, st0 = c->steps;
    const double t0 = c->toT, d1 = c->dt1, d2 = c->dt2;
    for (int k = 0; k < nsteps; ++k) {
      MarchArgs<T> a{};
      step_scalars<T>(c, a);
      double *r = &scal[((size_t)k * nm + q) * 4];
      r[0] = a.lam;
      r[1] = a.dt;
      r[2] = a.dtb;
      r[3] = a.amp;
      advance_time(c);
    }
    c->n = n0;
    c->steps = st0;
    c->toT = t0;
    c->dt1 = d1;
    c->dt2 = d2;
  }
  HIPCHK(hipMemcpy(b->d_scal, scal.data(), scal.size() * sizeof(double), hipMemcpyHostToDevice));
  int by = 0, pf = 0;
  batch_tiling(b, by, pf);
  MarchArgs<T> la;
  UpcontArgs<T> ul;
  const int pby = nsteps >= 2 ? batch_pair_rows(b) : 0;
  if (int rc = batch_records<T>(b, by, pf, la, ul, pby)) return rc;
  // the record parity flips once per step or pair (a pair leaves its result where one step would)
  int par = 0;
  for (int k = 0; k < nsteps; par ^= 1) {
    if (pby && k + 1 < nsteps && batch_pair_ok(b)) {
      if (int rc = batch_enqueue_pair<T>(b, la, ul, par, k, true)) return rc;
      k += 2;
    } else {
      if (int rc = batch_enqueue<T>(b, la, ul, par, k, true)) return rc;
      k += 1;
    }
  }
  b->steps += nsteps;
  return batch_join(b);
}

// Times batch steps (zeroed fields, no source, no sampling) for each rows x prefetch tiling and
// keeps the fastest; the members' fields are zeroed again afterwards.
template <typename T>
int batch_autotune_t(tcsem_batch *b, int nsteps, int *best_by, int *best_pf, int *best_s8, int *best_pairs,
                     double *best_ms) {
  const int nm = (int)b->m.size();
  if (int rc = batch_quiesce(b)) return rc;
  for (int q = 0; q < nm; ++q) b->active[q] = 1;
  if (int rc = batch_scal_reserve(b, 1)) return rc;
  const bool timing0 = b->timing;
  b->timing = false;
  std::vector<double> scal((size_t)nm * 4);
  for (int q = 0; q < nm; ++q) {
    scal[4 * q + 0] = 1.0;
    scal[4 * q + 1] = 1e-6;
    scal[4 * q + 2] = 1e-6;
    scal[4 * q + 3] = 0.0;
  }
  HIPCHK(hipMemcpy(b->d_scal, scal.data(), scal.size() * sizeof(double), hipMemcpyHostToDevice));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(TCSEM_EDEVICE, "event creation failed");
  }
  std::vector<int> rows_c = b->prec == 64 ? std::vector<int>{4, 8, 10, 12} : std::vector<int>{4, 8, 10, 16};
  double best = 1e300;
  int rc = TCSEM_OK;
  const bool any8 = std::all_of(b->m.begin(), b->m.end(), [](const tcsem_ctx *c) { return c->d_sig8 != nullptr; });
  const bool s80 = b->s8;
  *best_s8 = 0;
  for (int by : rows_c) {
    for (int pf = 1; pf <= 4 && rc == TCSEM_OK; ++pf)
      for (int s8 = any8 ? 1 : 0; s8 >= 0 && rc == TCSEM_OK; --s8) {  // the members' indices, and not
        b->s8 = s8 != 0;
        MarchArgs<T> la;
        UpcontArgs<T> ul;
        if ((rc = batch_records<T>(b, by, pf, la, ul))) break;
        int par = 0;
        if ((rc = batch_enqueue<T>(b, la, ul, par, 0, false))) break;  // first launches of this tiling
        par ^= 1;
        hipError_t e = hipStreamWaitEvent(b->stream, b->xev[1], 0);
        if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
        for (int k = 0; k < nsteps && e == hipSuccess && rc == TCSEM_OK; ++k, par ^= 1)
          rc = batch_enqueue<T>(b, la, ul, par, 0, false);
        if (rc) break;
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->xev[1], 0);
        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          rc = fail(TCSEM_EDEVICE, "batch autotune: %s", hipGetErrorString(e));
          break;
        }
        const double per = (double)ms / nsteps;
        if (per < best) {
          best = per;
          *best_by = supported_rows(b->prec, by);
          *best_pf = pf;
          *best_s8 = s8;
        }
      }
    if (rc) break;
  }
  b->s8 = s80;
  // two steps per pass with the one-step pick for the top launches: kept when 2% faster per step
  int pbest = 0;
  if (rc == TCSEM_OK && *best_by) {
    b->s8 = *best_s8 != 0;
    double pms = 1e300;
    if ((rc = batch_scal_reserve(b, 2)) == TCSEM_OK) {
      std::vector<double> scal2(scal);
      scal2.insert(scal2.end(), scal.begin(), scal.end());
      const hipError_t e = hipMemcpy(b->d_scal, scal2.data(), scal2.size() * sizeof(double), hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = fail(TCSEM_EDEVICE, "batch autotune: %s", hipGetErrorString(e));
    }
    const std::vector<int> prow = b->prec == 64 ? std::vector<int>{12, 11, 10, 8} : std::vector<int>{12, 14, 16, 10, 8};
    const int npair = std::max(1, nsteps / 2);
    // two interleaved rounds over the pass heights; each keeps its faster time
    std::vector<double> pm(prow.size(), 1e300);
    for (int round = 0; round < 2 * (int)prow.size(); ++round) {
      if (rc) break;
      const size_t ir = round % prow.size();
      const int r = prow[ir];
      MarchArgs<T> la;
      UpcontArgs<T> ul;
      if ((rc = batch_records<T>(b, *best_by, *best_pf, la, ul, r))) break;
      int par = 0;
      if ((rc = batch_enqueue_pair<T>(b, la, ul, par, 0, false))) break;  // first launches of this shape
      par ^= 1;
      hipError_t e = hipStreamWaitEvent(b->stream, b->xev[1], 0);
      if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
      for (int k = 0; k < npair && e == hipSuccess && rc == TCSEM_OK; ++k, par ^= 1)
        rc = batch_enqueue_pair<T>(b, la, ul, par, 0, false);
      if (rc) break;
      if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->xev[1], 0);
      if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(TCSEM_EDEVICE, "batch autotune: %s", hipGetErrorString(e));
        break;
      }
      pm[ir] = std::min(pm[ir], (double)ms / (2.0 * npair));
    }
    for (size_t ir = 0; ir < prow.size() && rc == TCSEM_OK; ++ir)
      if (pm[ir] < pms) {
        pms = pm[ir];
        pbest = prow[ir];
      }
    b->s8 = s80;
    if (rc == TCSEM_OK && pbest && pms < 0.98 * best) {
      best = pms;
    } else {
      pbest = 0;
    }
  }
  *best_pairs = pbest;
  b->pair_by = 0;  // the records are rebuilt by the next step call
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->timing = timing0;
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(b->stream2));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (tcsem_ctx *c : b->m)
    if ((rc = tcsem_reset_fields(c, 0, 1))) return rc;
  *best_ms = best;
  return TCSEM_OK;
}

}  // namespace

extern "C" {

const char *tcsem_version(void) { return "tcsem-mi355x 0.1 (gfx950)"; }
const char *tcsem_last_error(void) { return g_err.c_str(); }

int tcsem_read_setup(const char *path, tcsem_setup *o) {  // ModelingSetup.cpp:43-61
  if (!path || !o) return fail(TCSEM_EINVAL, "null argument");
  FILE *f = fopen(path, "r");
  if (!f) return fail(TCSEM_EIO, "cannot open %s", path);
  int n1 = 0, ok = 1;
  ok &= fscanf(f, "%lf %lf\n", &o->sfa1, &o->sfa2) == 2;
  ok &= fscanf(f, "%d %lf\n", &n1, &o->sigma_t) == 2;
  ok &= fscanf(f, "%lf %lf %lf\n", &o->x0, &o->y0, &o->z0) == 3;
  ok &= fscanf(f, "%lf %lf %lf\n", &o->dxmin, &o->dymin, &o->dzmin) == 3;
  ok &= fscanf(f, "%lf %lf\n", &o->arg1, &o->arg2) == 2;
  ok &= fscanf(f, "%d %d %d %d %d\n", &o->ALx2, &o->ALx1, &o->Nx, &o->ARx1, &o->ARx2) == 5;
  ok &= fscanf(f, "%d %d %d\n", &o->Ay2, &o->Ay1, &o->Ny) == 3;
  ok &= fscanf(f, "%d %d %d\n", &o->Az2, &o->Az1, &o->Nz) == 3;
  fclose(f);
  o->n_src = n1;
  if (!ok) return fail(TCSEM_EIO, "%s: malformed setup file", path);
  return TCSEM_OK;
}

int tcsem_read_cond(const char *path, size_t ncell, double *out) {
  if (!path || !out) return fail(TCSEM_EINVAL, "null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(TCSEM_EIO, "cannot open %s", path);
  const size_t got = fread(out, sizeof(double), ncell, f);
  fclose(f);
  if (got != ncell) return fail(TCSEM_EIO, "%s: %zu of %zu values", path, got, ncell);
  return TCSEM_OK;
}

int tcsem_read_receivers(const char *path, tcsem_rxtime *rt, int *nrx, double *xyz, int cap) {
  if (!path || !rt || !nrx) return fail(TCSEM_EINVAL, "null argument");
  FILE *f = fopen(path, "r");
  if (!f) return fail(TCSEM_EIO, "cannot open %s", path);
  int ok = 1;
  ok &= fscanf(f, "%lf %lf\n", &rt->T0, &rt->Tn) == 2;
  ok &= fscanf(f, "%d %d\n", &rt->Nt, &rt->is_log) == 2;
  ok &= fscanf(f, "%d %d %d %d %d %d\n", &rt->is_comp[0], &rt->is_comp[1], &rt->is_comp[2], &rt->is_comp[3],
               &rt->is_comp[4], &rt->is_comp[5]) == 6;
  ok &= fscanf(f, "%d\n", nrx) == 1;
  if (!ok || *nrx < 0) {
    fclose(f);
    return fail(TCSEM_EIO, "%s: malformed header", path);
  }
  if (xyz) {
    if (cap < *nrx) {
      fclose(f);
      return fail(TCSEM_ESIZE, "receiver buffer holds %d of %d", cap, *nrx);
    }
    for (int r = 0; r < *nrx; ++r) {
      int id;
      double x, y, z;
      if (fscanf(f, "%d %lf %lf %lf\n", &id, &x, &y, &z) != 4) {
        fclose(f);
        return fail(TCSEM_EIO, "%s: only %d of %d receivers", path, r, *nrx);
      }
      if (id < 1 || id > *nrx) {
        fclose(f);
        return fail(TCSEM_EIO, "%s: receiver id %d out of range", path, id);
      }
      xyz[3 * (id - 1)] = x;
      xyz[3 * (id - 1) + 1] = y;
      xyz[3 * (id - 1) + 2] = z;
    }
  }
  fclose(f);
  return TCSEM_OK;
}

int tcsem_read_transmitters(const char *path, tcsem_tx *out, int cap, int *ntx, int *ids, int ids_cap,
                            int *nids) {
  if (!path || !ntx) return fail(TCSEM_EINVAL, "null argument");
  FILE *f = fopen(path, "r");
  if (!f) return fail(TCSEM_EIO, "cannot open %s", path);
  if (fscanf(f, "%d\n", ntx) != 1 || *ntx < 0) {
    fclose(f);
    return fail(TCSEM_EIO, "%s: malformed header", path);
  }
  int used = 0, rc = TCSEM_OK;
  for (int r = 0; r < *ntx && rc == TCSEM_OK; ++r) {
    tcsem_tx t{};
    int itx = 0, ort = 0;
    if (fscanf(f, "%d %d", &itx, &ort) != 2) {
      rc = fail(TCSEM_EIO, "%s: row %d", path, r + 1);
      break;
    }
    t.id = itx;
    t.ort = ort;
    int ok;
    if (ort < 4) ok = fscanf(f, " %lf %lf %lf %lf %d", &t.bipolen, &t.x, &t.y, &t.z, &t.nr) == 5;
    else if (ort == 4)
      ok = fscanf(f, " %lf %lf %lf %lf %lf %lf %d", &t.bipolen, &t.x, &t.y, &t.z, &t.azi_deg, &t.dip_deg, &t.nr) == 7;
    else if (ort == 5)
      ok = fscanf(f, " %lf %lf %lf %lf %lf %lf", &t.bipolen, &t.x, &t.y, &t.z, &t.azi_deg, &t.dip_deg) == 6 &&
           fscanf(f, " %lf %lf %lf %lf %lf %lf %d", &t.bipolen1, &t.x1, &t.y1, &t.z1, &t.azi1_deg, &t.dip1_deg,
                  &t.nr) == 7;
    else {
      rc = fail(TCSEM_EINVAL, "%s: illegal transmitter type %d", path, ort);
      break;
    }
    if (!ok || t.nr < 0) {
      rc = fail(TCSEM_EIO, "%s: row %d", path, r + 1);
      break;
    }
    int *dst = (ids && used + t.nr <= ids_cap) ? ids + used : nullptr;
    for (int q = 0; q < t.nr; ++q) {
      int v;
      if (fscanf(f, " %d", &v) != 1) {
        rc = fail(TCSEM_EIO, "%s: row %d receiver list", path, r + 1);
        break;
      }
      if (dst) dst[q] = v;
    }
    if (out) {
      if (itx < 1 || itx > cap) {
        rc = fail(TCSEM_EIO, "%s: transmitter id %d out of range", path, itx);
        break;
      }
      if (!dst) {
        rc = fail(TCSEM_ESIZE, "receiver id buffer too small");
        break;
      }
      t.rx_ids = dst;
      out[itx - 1] = t;
    }
    used += t.nr;
  }
  fclose(f);
  if (nids) *nids = used;
  return rc;
}

int tcsem_write_bout(const char *path, int nr, int ncomp, int nt, const double *data) {
  if (!path || !data || nr < 0 || ncomp < 0 || nt < 0) return fail(TCSEM_EINVAL, "bad argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(TCSEM_EIO, "cannot open %s", path);
  const size_t n = (size_t)nr * ncomp * nt;
  bool ok = fwrite(&nr, sizeof(int), 1, f) == 1 && fwrite(&ncomp, sizeof(int), 1, f) == 1 &&
            fwrite(&nt, sizeof(int), 1, f) == 1 && fwrite(data, sizeof(double), n, f) == n;
  ok = (fclose(f) == 0) && ok;
  return ok ? TCSEM_OK : fail(TCSEM_EIO, "write to %s failed", path);
}

int tcsem_read_bout(const char *path, int *nr, int *ncomp, int *nt, double *data, size_t cap) {
  if (!path || !nr || !ncomp || !nt) return fail(TCSEM_EINVAL, "null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(TCSEM_EIO, "cannot open %s", path);
  bool ok = fread(nr, sizeof(int), 1, f) == 1 && fread(ncomp, sizeof(int), 1, f) == 1 &&
            fread(nt, sizeof(int), 1, f) == 1;
  if (ok && data) {
    const size_t n = (size_t)(*nr) * (*ncomp) * (*nt);
    if (n > cap) {
      fclose(f);
      return fail(TCSEM_ESIZE, "buffer holds %zu of %zu values", cap, n);
    }
    ok = fread(data, sizeof(double), n, f) == n;
  }
  fclose(f);
  return ok ? TCSEM_OK : fail(TCSEM_EIO, "%s: truncated", path);
}

int tcsem_grid_info_of(const tcsem_setup *s, tcsem_grid_info *out) {
  if (!s || !out) return fail(TCSEM_EINVAL, "null argument");
  HostGrid g;
  const int rc = build_grid(*s, g);
  if (rc) return rc;
  *out = g.info;
  return TCSEM_OK;
}

int tcsem_grid_arrays(const tcsem_setup *s, double *dx, double *dy, double *dz, double *X, double *Y, double *Z,
                      double *Xnew, double *Ynew) {
  if (!s) return fail(TCSEM_EINVAL, "null argument");
  HostGrid g;
  const int rc = build_grid(*s, g);
  if (rc) return rc;
  build_maps(*s, g);
  auto cp = [](double *d, const std::vector<double> &v) {
    if (d) std::memcpy(d, v.data(), v.size() * sizeof(double));
  };
  cp(dx, g.dx);
  cp(dy, g.dy);
  cp(dz, g.dz);
  cp(X, g.X);
  cp(Y, g.Y);
  cp(Z, g.Z);
  cp(Xnew, g.Xn);
  cp(Ynew, g.Yn);
  return TCSEM_OK;
}

int tcsem_grid_maps(const tcsem_setup *s, int *a, int *b, int *c, int *d, int *e, int *f) {
  if (!s) return fail(TCSEM_EINVAL, "null argument");
  HostGrid g;
  const int rc = build_grid(*s, g);
  if (rc) return rc;
  build_maps(*s, g);
  auto cp = [](int *dd, const std::vector<int> &v) {
    if (dd) std::memcpy(dd, v.data(), v.size() * sizeof(int));
  };
  cp(a, g.xN2o);
  cp(b, g.yN2o);
  cp(c, g.xo2N);
  cp(d, g.yo2N);
  cp(e, g.xBx2N);
  cp(f, g.yBy2N);
  return TCSEM_OK;
}

int tcsem_waveform(int ns, double sigma_t, double *t, double *dt, double *wave) {
  if (!t || !dt || !wave) return fail(TCSEM_EINVAL, "null argument");
  return waveform(ns, sigma_t, t, dt, wave);
}

int tcsem_create(const tcsem_setup *s, const double *cond, int precision, tcsem_ctx **out) {
  if (!s || !cond || !out) return fail(TCSEM_EINVAL, "null argument");
  if (precision != TCSEM_FP64 && precision != TCSEM_FP32) return fail(TCSEM_EINVAL, "precision must be 32 or 64");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TCSEM_EDEVICE, "no HIP device");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(TCSEM_EDEVICE, "device %s is not gfx950 (MI355X); this build targets gfx950 only", prop.gcnArchName);
  // the allocation-free entry point's work planes on this device (looked up once per device)
  if (hipError_t e = upcont_scratch_init(dev); e != hipSuccess)
    return fail(TCSEM_EDEVICE, "scratch address on device %d: %s", dev, hipGetErrorString(e));
  tcsem_ctx *c = new tcsem_ctx();
  c->dev = dev;
  c->prec = precision;
  c->esz = precision == TCSEM_FP64 ? 8 : 4;
  c->setup = *s;
  int rc = build_grid(*s, c->g);
  if (rc) {
    delete c;
    return rc;
  }
  build_maps(*s, c->g);
  const tcsem_grid_info &I = c->g.info;
  const int L = I.L, M = I.M, N = I.N;
  if (s->n_src / 2 < 1 || !(s->sigma_t > 0)) {
    delete c;
    return fail(TCSEM_EINVAL, "source discretisation must be >= 2 and sigma_t > 0");
  }
  const size_t ncell = (size_t)L * M * N;
  c->cond.assign(cond, cond + ncell);
  for (size_t q = 0; q < ncell; ++q)
    if (!(c->cond[q] >= 0.0) || !std::isfinite(c->cond[q])) {
      delete c;
      return fail(TCSEM_EINVAL, "conductivity must be finite and >= 0 (cell %zu)", q);
    }
  c->xm = kLineBytes / (int)c->esz;
  c->PX = ((L + 2 + c->xm) + 63) / 64 * 64;
  c->PY = M + 3;
  c->PS = (long long)c->PX * c->PY;
  c->fstride = (long long)(N + 1) * c->PS;
  c->pstride = c->PS;
  c->wstride = (long long)I.L1 * I.M1;
  choose_tiling(c, 1);
  // air-plane operators
  c->n1 = c->g.AL + kNPad;
  c->n2 = c->g.AR + kNPad;
  c->n3 = c->g.AY + kNPad;
  {
    const HostGrid &g = c->g;
    const int aL = I.AnLx_New, aR = I.AnRx_New, aY = I.Any_New, Nx = s->Nx, Ny = s->Ny;
    const int ox = Nx + g.AL - kNPad, oy = Ny + g.AY - kNPad;
    std::vector<int> idr(std::max(std::max(aL, aR), aY));
    for (int p = 0; p < aL; ++p) idr[p] = g.xN2o[p];
    spline_weights(c->n1, g.dx.data(), g.X.data(), aL, g.Xn.data(), idr.data(), c->wl);
    for (int p = 0; p < aR; ++p) idr[p] = g.xN2o[aL + Nx + p] - ox;
    spline_weights(c->n2, g.dx.data() + ox, g.X.data() + ox, aR, g.Xn.data() + aL + Nx, idr.data(), c->wr);
    for (int p = 0; p < aY; ++p) idr[p] = g.yN2o[p];
    spline_weights(c->n3, g.dy.data(), g.Y.data(), aY, g.Yn.data(), idr.data(), c->wf);
    for (int p = 0; p < aY; ++p) idr[p] = g.yN2o[aY + Ny + p] - oy;
    spline_weights(c->n3, g.dy.data() + oy, g.Y.data() + oy, aY, g.Yn.data() + aY + Ny, idr.data(), c->wb);
    // wavenumber gains on the half spectrum (UpContinuation_gpu.cu:113-164)
    const int L1 = I.L1, M1 = I.M1, H = L1 / 2 + 1;
    c->gx.assign((size_t)M1 * H, 0.0);
    c->gy.assign((size_t)M1 * H, 0.0);
    const double dzdx = s->dzmin / s->dxmin, dzdy = s->dzmin / s->dymin;
    const double rxy = s->dxmin / s->dymin, ryx = s->dymin / s->dxmin;
    for (int iy = 0; iy < M1; ++iy)
      for (int ix = 0; ix < H; ++ix) {
        const double v1 = (iy < M1 / 2 + 1) ? (1.0 * kPi * iy) / M1 * dzdy : -(1.0 * kPi * (M1 - iy)) / M1 * dzdy;
        const double u1 = (1.0 * kPi * ix) / L1 * dzdx;
        double gxv = 1.0, gyv = 1.0;
        if (!(ix == 0 && iy == 0)) {
          const double rr = u1 * u1 + v1 * v1;
          const double upcc = std::exp(-std::sqrt(rr)) * (1.0 / std::sqrt(rr));
          gxv = rxy * u1 * upcc;
          gyv = ryx * v1 * upcc;
        }
        c->gx[(size_t)iy * H + ix] = gxv;
        c->gy[(size_t)iy * H + ix] = gyv;
      }
    make_plan(L1, c->plx);
    make_plan(M1, c->ply);
    // The tiled chain needs both working lengths <= 2048 (one wavefront per line, two LDS buffers)
    // and stages an irregular row pair (L) or column (M) in one line buffer of m + 2 entries.
    const char *chain = std::getenv("TCSEM_UPCONT_CHAIN");
    const bool want_lines = chain && std::strcmp(chain, "lines") == 0;
    c->up_tiled = !want_lines && c->plx.m <= 2048 && c->ply.m <= 2048 && L <= c->plx.m + 2 && M <= c->ply.m + 2;
    if (c->up_tiled) {
      // pitches are whole 128-byte lines and avoid multiples of 1 KiB (HBM channel spread)
      auto pitch = [](int n, int per_line) {
        int p = (n + per_line - 1) / per_line * per_line;
        if (p % (8 * per_line) == 0) p += per_line;
        return p;
      };
      auto slot_stride = [](long long n) {
        n = (n + 31) / 32 * 32;
        if (n % 2048 == 0) n += 32;
        return n;
      };
      // in entries of the field type: 128 bytes hold 128/esz reals, half as many complex
      const int per = kLineBytes / (int)c->esz;
      c->MA = pitch(M, per / 2);    // complex
      c->MG = pitch(M1, per / 4);   // (Gx, Gy) pairs of complex
      c->XR = pitch(L1, per);       // reals
      c->sA = slot_stride(H * (long long)c->MA);       // complex entries
      c->sG = slot_stride(2LL * H * c->MG);            // complex entries
      c->sX = slot_stride((long long)M1 * c->XR);      // reals
      c->gxT.assign((size_t)H * M1, 0.0);
      c->gyT.assign((size_t)H * M1, 0.0);
      for (int ky = 0; ky < M1; ++ky)
        for (int kx = 0; kx < H; ++kx) {
          double gxv = c->gx[(size_t)ky * H + kx], gyv = c->gy[(size_t)ky * H + kx];
          if (kx == 0 && ky == 0) gxv = gyv = 0.0;
          if (L1 % 2 == 0 && 2 * kx == L1) gxv = 0.0;
          if (M1 % 2 == 0 && 2 * ky == M1 && (kx == 0 || (L1 % 2 == 0 && 2 * kx == L1))) gyv = 0.0;
          c->gxT[(size_t)kx * M1 + ky] = gxv;
          c->gyT[(size_t)kx * M1 + ky] = gyv;
        }
    }
  }
  const size_t lds_max = 160 * 1024 - 1024;
  // the march's masked stores go through a buffer resource per slot field that must stay below
  // 2 GiB (out-of-range offsets are its "drop" mask)
  if ((unsigned long long)c->fstride * c->esz >= (1ull << 31)) {
    delete c;
    return fail(TCSEM_ESIZE, "one field of %d x %d x %d cells exceeds 2 GiB", L, M, N);
  }
  // the march keeps 3 N z-coefficients in LDS beside at most 64 KiB of tiles
  if ((size_t)3 * N * c->esz + 64 * 1024 > lds_max) {
    delete c;
    return fail(TCSEM_ESIZE, "%d levels exceed the march's LDS budget", N);
  }
  if ((size_t)c->ply.m * 16 + ((size_t)M + 1) / 2 * 16 > lds_max || (size_t)c->plx.m * 16 > lds_max) {
    const int mx = c->plx.m, my = c->ply.m;
    delete c;
    return fail(TCSEM_ESIZE, "air plane %d x %d needs FFT lengths %d/%d beyond the LDS budget", I.L1, I.M1, mx,
                my);
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return fail(TCSEM_EDEVICE, "stream creation failed");
  }
  if (hipStreamCreate(&c->stream2) != hipSuccess) {
    tcsem_destroy(c);
    return fail(TCSEM_EDEVICE, "stream creation failed");
  }
  for (auto &e : c->xev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      tcsem_destroy(c);
      return fail(TCSEM_EDEVICE, "event creation failed");
    }
  // both ordering events start recorded, so the first step's waits are satisfied
  if (hipEventRecord(c->xev[0], c->stream) != hipSuccess || hipEventRecord(c->xev[1], c->stream2) != hipSuccess) {
    tcsem_destroy(c);
    return fail(TCSEM_EDEVICE, "event record failed");
  }
  // device constants
  auto bail = [&](int code) {
    tcsem_destroy(c);
    return code;
  };
  {
    std::vector<double> pad((size_t)N * c->PS, 0.0);
    for (int k = 0; k < N; ++k)
      for (int j = 0; j < M; ++j)
        for (int i = 0; i < L; ++i)
          pad[(size_t)k * c->PS + (size_t)(j + 1) * c->PX + (i + c->xm)] = c->cond[((size_t)k * M + j) * L + i];
    if (hipMalloc(&c->d_sig, pad.size() * c->esz) != hipSuccess) return bail(fail(TCSEM_ENOMEM, "sigma allocation"));
    if ((rc = upload_double_as(c, c->d_sig, pad))) return bail(rc);
    // the 8-bit index: distinct values as the field precision holds them, 0 first
    const char *s8 = std::getenv("TCSEM_SIGMA8");
    if (!(s8 && std::strcmp(s8, "0") == 0)) {
      auto bits_t = [&](double v) {  // the value's bit pattern in the field precision
        std::uint64_t b = 0;
        if (precision == TCSEM_FP64) {
          std::memcpy(&b, &v, 8);
        } else {
          const float f = (float)v;
          std::uint32_t u;
          std::memcpy(&u, &f, 4);
          b = u;
        }
        return b;
      };
      std::vector<double> lut{0.0};
      std::unordered_map<std::uint64_t, unsigned char> at{{bits_t(0.0), 0}};
      std::vector<unsigned char> idx(pad.size(), 0);
      bool fits = true;
      std::uint64_t last = bits_t(0.0);
      unsigned char last_e = 0;
      for (size_t q = 0; q < pad.size(); ++q) {
        const std::uint64_t b = bits_t(pad[q]);
        if (b != last) {
          auto it = at.find(b);
          if (it == at.end()) {
            if (lut.size() == 256) {
              fits = false;
              break;
            }
            it = at.emplace(b, (unsigned char)lut.size()).first;
            lut.push_back(pad[q]);
          }
          last = b;
          last_e = it->second;
        }
        idx[q] = last_e;
      }
      if (fits) {
        c->sig8_n = (int)lut.size();
        lut.resize(256, 0.0);
        if (hipMalloc((void **)&c->d_sig8, idx.size()) != hipSuccess || hipMalloc(&c->d_slut, 256 * c->esz) != hipSuccess)
          return bail(fail(TCSEM_ENOMEM, "sigma index allocation"));
        if (hipMemcpy(c->d_sig8, idx.data(), idx.size(), hipMemcpyHostToDevice) != hipSuccess)
          return bail(fail(TCSEM_EDEVICE, "sigma index upload"));
        if ((rc = upload_double_as(c, c->d_slut, lut))) return bail(rc);
      }
    }
  }
  {
    const HostGrid &g = c->g;
    std::vector<double> rd(2 * (L + 3) + 2 * (M + 3) + 3 * N, 0.0);
    double *rdx = rd.data(), *rdx2 = rdx + (L + 3), *rdy = rdx2 + (L + 3), *rdy2 = rdy + (M + 3),
           *dzt = rdy2 + (M + 3), *rdz = dzt + N, *rdz2 = rdz + N;
    for (int i = 0; i < L; ++i) rdx[i + 1] = 1.0 / g.dx[i];
    for (int i = 1; i < L; ++i) rdx2[i + 1] = 1.0 / (g.dx[i - 1] + g.dx[i]);
    for (int j = 0; j < M; ++j) rdy[j + 1] = 1.0 / g.dy[j];
    for (int j = 1; j < M; ++j) rdy2[j + 1] = 1.0 / (g.dy[j - 1] + g.dy[j]);
    for (int k = 0; k < N; ++k) {
      dzt[k] = g.dz[k];
      rdz[k] = 1.0 / g.dz[k];
      rdz2[k] = k ? 1.0 / (g.dz[k] + g.dz[k - 1]) : 1.0 / (2.0 * g.dz[0]);
    }
    if (hipMalloc(&c->d_rd, rd.size() * c->esz) != hipSuccess) return bail(fail(TCSEM_ENOMEM, "grid allocation"));
    if ((rc = upload_double_as(c, c->d_rd, rd))) return bail(rc);
    std::vector<double> gd;
    for (auto *v : {&g.X, &g.Y, &g.Z, &g.dx, &g.dy, &g.dz, &g.Xn, &g.Yn}) gd.insert(gd.end(), v->begin(), v->end());
    if (hipMalloc((void **)&c->d_grid, gd.size() * 8) != hipSuccess) return bail(fail(TCSEM_ENOMEM, "grid allocation"));
    if (hipMemcpy(c->d_grid, gd.data(), gd.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(TCSEM_EDEVICE, "grid upload"));
    std::vector<int> mp;
    for (auto *v : {&g.xBx2N, &g.yo2N, &g.xo2N, &g.yBy2N}) mp.insert(mp.end(), v->begin(), v->end());
    // the regular rows the bilinear step reads (rows JJ-1 and JJ of every BxAir and ByAir row,
    // UpContinuation_gpu.cu:42-110): the tiled chain's row inverse computes only these
    {
      const int M1 = g.info.M1;
      std::vector<char> need(M1, 0);
      auto mark = [&](int jj) {
        jj = std::min(std::max(jj, 1), M1 - 1);
        need[jj - 1] = need[jj] = 1;
      };
      for (int iy = 0; iy < g.info.M; ++iy) mark(g.yo2N[iy]);
      for (int iy = 0; iy <= g.info.M; ++iy) mark(g.yBy2N[iy]);
      c->nyr = 0;
      for (int y = 0; y < M1; ++y)
        if (need[y]) {
          mp.push_back(y);
          ++c->nyr;
        }
    }
    if (hipMalloc((void **)&c->d_maps, mp.size() * 4) != hipSuccess) return bail(fail(TCSEM_ENOMEM, "map allocation"));
    if (hipMemcpy(c->d_maps, mp.data(), mp.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(TCSEM_EDEVICE, "map upload"));
    std::vector<double> up;
    // x-window weights go tap-major (n x columns) so that lanes over columns read them coalesced
    auto tap_major = [](const std::vector<double> &w, int n) {
      const size_t rows = n ? w.size() / n : 0;
      std::vector<double> t(w.size());
      for (size_t r = 0; r < rows; ++r)
        for (int q = 0; q < n; ++q) t[(size_t)q * rows + r] = w[r * n + q];
      return t;
    };
    const std::vector<double> wlt = tap_major(c->wl, c->n1), wrt = tap_major(c->wr, c->n2);
    for (const std::vector<double> *v : {&wlt, &wrt, (const std::vector<double> *)&c->wf, (const std::vector<double> *)&c->wb,
                                         (const std::vector<double> *)&c->gx, (const std::vector<double> *)&c->gy})
      up.insert(up.end(), v->begin(), v->end());
    for (auto *P : {&c->plx, &c->ply})
      for (auto *v : {&P->tw, &P->chirp, &P->bhat}) up.insert(up.end(), v->begin(), v->end());
    up.insert(up.end(), c->gxT.begin(), c->gxT.end());
    up.insert(up.end(), c->gyT.begin(), c->gyT.end());
    if (hipMalloc((void **)&c->d_up, up.size() * 8) != hipSuccess) return bail(fail(TCSEM_ENOMEM, "operator allocation"));
    if (hipMemcpy(c->d_up, up.data(), up.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(TCSEM_EDEVICE, "operator upload"));
  }
  if ((rc = alloc_slots(c, 1))) return bail(rc);
  c->nbatch = 1;
  c->up64 = upcont_args<double>(c);
  c->up32 = upcont_args<float>(c);
  // empty per-slot source/receiver tables so single-slot kernels can run before a batch starts
  {
    std::vector<int> z(2, 0);
    if (hipMalloc((void **)&c->d_srcidx, 2 * sizeof(int)) != hipSuccess) return bail(fail(TCSEM_ENOMEM, "alloc"));
    if (hipMemcpy(c->d_srcidx, z.data(), 2 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(TCSEM_EDEVICE, "upload"));
  }
  // graph replays by default where the host's launches bound a step (grids under 2^21 cells)
  if ((long long)L * M * N < kGraphAutoCells) {
    if (hipMalloc((void **)&c->d_gs, (size_t)4 * kGraphMaxSteps * sizeof(double)) != hipSuccess)
      return bail(fail(TCSEM_ENOMEM, "graph scalar allocation"));
    c->graph_k = kGraphAutoSteps;
  }
  *out = c;
  return TCSEM_OK;
}

void tcsem_destroy(tcsem_ctx *c) {
  if (!c) return;
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_run(c);
  (void)hipFree(c->d_fields);
  (void)hipFree(c->d_sig);
  (void)hipFree(c->d_sig8);
  (void)hipFree(c->d_slut);
  (void)hipFree(c->d_planes);
  (void)hipFree(c->d_zs);
  (void)hipFree(c->d_ib);
  (void)hipFree(c->d_mid);
  (void)hipFree(c->d_gs);
  (void)hipFree(c->d_rd);
  (void)hipFree(c->d_grid);
  (void)hipFree(c->d_maps);
  (void)hipFree(c->d_up);
  (void)hipFree(c->d_work);
  (void)hipFree(c->d_scratch);
  (void)hipFree(c->d_rx);
  (void)hipFree(c->d_rxi);
  for (auto &e : c->xev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->evpool)
    if (e) (void)hipEventDestroy(e);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int tcsem_info(const tcsem_ctx *c, tcsem_grid_info *out) {
  if (!c || !out) return fail(TCSEM_EINVAL, "null argument");
  *out = c->g.info;
  return TCSEM_OK;
}

int tcsem_set_receivers(tcsem_ctx *c, const tcsem_rxtime *rt, int nrx, const double *xyz) {
  if (check_ctx(c) || !rt || (nrx > 0 && !xyz) || nrx < 0) return fail(TCSEM_EINVAL, "bad receiver arguments");
  if (rt->Nt < 2) return fail(TCSEM_EINVAL, "need at least 2 sample times");
  if (rt->is_log && !(rt->T0 > 0)) return fail(TCSEM_EINVAL, "log-spaced sampling needs T0 > 0");
  const tcsem_grid_info &I = c->g.info;
  const int L = I.L, M = I.M, K = I.N;
  std::vector<double> x(nrx), y(nrx), z(nrx);
  std::vector<int> ix(nrx), iy(nrx), iz(nrx);
  for (int r = 0; r < nrx; ++r) {  // LoadRxTxPos.cpp:93-121
    x[r] = xyz[3 * r];
    y[r] = xyz[3 * r + 1];
    z[r] = xyz[3 * r + 2];
    ix[r] = search_insert(c->g.X.data(), L, x[r]);
    iy[r] = search_insert(c->g.Y.data(), M, y[r]);
    iz[r] = search_insert(c->g.Z.data(), K, z[r]);
    // the 4-point stencils need two nodes on each side (the reference reads outside the arrays)
    if (ix[r] < 2 || ix[r] > L - 3) return fail(TCSEM_EDOMAIN, "receiver %d too close to the x boundary", r + 1);
    if (iy[r] < 2 || iy[r] > M - 3) return fail(TCSEM_EDOMAIN, "receiver %d too close to the y boundary", r + 1);
    if (iz[r] > K - 3) return fail(TCSEM_EDOMAIN, "receiver %d too close to the bottom", r + 1);
  }
  free_run(c);
  c->rt = *rt;
  c->nrx = nrx;
  c->rxx = x;
  c->rxy = y;
  c->rxz = z;
  c->rix = ix;
  c->riy = iy;
  c->riz = iz;
  c->ncomp = 0;
  c->nefcmp = 0;
  for (int q = 0; q < 6; ++q)
    if (rt->is_comp[q]) c->comps[c->ncomp++] = q;
  for (int q = 0; q < 3; ++q) c->nefcmp += rt->is_comp[q] != 0;
  sample_times(*rt, c->times);
  (void)hipFree(c->d_rx);
  (void)hipFree(c->d_rxi);
  c->d_rx = nullptr;
  c->d_rxi = nullptr;
  const int nn = std::max(nrx, 1);
  std::vector<double> hx(3 * nn, 0.0);
  std::vector<int> hi(3 * nn, 0);
  for (int r = 0; r < nrx; ++r) {
    hx[r] = x[r];
    hx[nn + r] = y[r];
    hx[2 * nn + r] = z[r];
    hi[r] = ix[r];
    hi[nn + r] = iy[r];
    hi[2 * nn + r] = iz[r];
  }
  if (hipMalloc((void **)&c->d_rx, hx.size() * 8) != hipSuccess || hipMalloc((void **)&c->d_rxi, hi.size() * 4) != hipSuccess)
    return fail(TCSEM_ENOMEM, "receiver allocation");
  HIPCHK(hipMemcpy(c->d_rx, hx.data(), hx.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_rxi, hi.data(), hi.size() * 4, hipMemcpyHostToDevice));
  // receiver tables are addressed with stride nn; keep nrx for id checks
  c->nrx = nrx;
  c->have_rx = true;
  // re-locate transmitters against the new receiver list
  c->tx.clear();
  return TCSEM_OK;
}

int tcsem_set_transmitters(tcsem_ctx *c, const tcsem_tx *tx, int ntx) {
  if (check_ctx(c) || (ntx > 0 && !tx) || ntx < 0) return fail(TCSEM_EINVAL, "bad transmitter arguments");
  if (!c->have_rx) return fail(TCSEM_ESTATE, "set receivers before transmitters");
  std::vector<TxLoc> loc(ntx);
  for (int q = 0; q < ntx; ++q) {
    if (tx[q].nr > 0 && !tx[q].rx_ids) return fail(TCSEM_EINVAL, "transmitter %d has no receiver list", tx[q].id);
    const int rc = locate_tx(c, tx[q], loc[q]);
    if (rc) return rc;
    source_cells(c, loc[q]);
  }
  free_run(c);
  c->tx = std::move(loc);
  return TCSEM_OK;
}

int tcsem_reserve_batch(tcsem_ctx *c, int nbatch) {
  if (check_ctx(c) || nbatch < 1) return fail(TCSEM_EINVAL, "bad batch size");
  const int rc = alloc_slots(c, nbatch);
  if (rc) return rc;
  c->up64 = upcont_args<double>(c);
  c->up32 = upcont_args<float>(c);
  return TCSEM_OK;
}

int tcsem_reset_fields(tcsem_ctx *c, int first, int n) {
  if (check_ctx(c) || first < 0 || n < 0 || first + n > c->nslots) return fail(TCSEM_EINVAL, "bad slot range");
  c->ib_valid = false;
  const size_t fb = (size_t)c->fstride * c->esz, pb = (size_t)c->pstride * c->esz;
  for (int f = 0; f < 10; ++f)
    HIPCHK(launch_fill_zero(c->d_fields + ((size_t)f * c->nslots + first) * fb, fb * n, c->stream));
  for (int f = 0; f < 7; ++f)
    HIPCHK(launch_fill_zero(c->d_planes + ((size_t)f * c->nslots + first) * pb, pb * n, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return TCSEM_OK;
}

// Lists of the source pass that follows the volume march (levels 1 .. N-1): every E cell of a
// slot's table once, with the weights of coincident entries summed in table order (the order of
// the per-cell loop of ElectricalFiled.cu:347-475), then every Bx/By cell whose curl reads one of
// them (Bderivate.cu:79-200), sorted by (component, j, i, k) and cut into one run per component
// and column.  Cells outside the update masks are dropped, as the march drops them; level-0 cells
// stay with the surface kernel.
static int build_fix_lists(tcsem_ctx *c, int nbatch) {
  const int L = c->g.info.L, M = c->g.info.M, N = c->g.info.N;
  auto e_ok = [&](const SrcCell &s) {
    if (s.k < 1 || s.k > N - 1 || s.i < 0 || s.j < 0) return false;
    if (s.comp == 0) return s.i < L && s.j >= 1 && s.j < M;
    if (s.comp == 1) return s.i >= 1 && s.i < L && s.j < M;
    return s.i >= 1 && s.i < L && s.j >= 1 && s.j < M;
  };
  auto b_ok = [&](int comp, int k, int j, int i) {
    if (k < 1 || k > N - 1 || i < 0 || j < 0) return false;
    if (comp == 3) return i >= 1 && i < L && j < M;
    return i < L && j >= 1 && j < M;
  };
  std::vector<SrcCell> ecells, bcells;
  std::vector<int2> runs;
  std::vector<int> idx(4 * nbatch, 0);
  for (int b = 0; b < nbatch; ++b) {
    idx[b] = (int)ecells.size();
    idx[2 * nbatch + b] = (int)runs.size();
    if (!c->tx.empty()) {
      const TxLoc &t = c->tx[c->tx_first + b];
      const size_t e0 = ecells.size(), b0 = bcells.size();
      for (const SrcCell &s : t.cells) {
        if (!e_ok(s)) continue;
        bool merged = false;
        for (size_t q = e0; q < ecells.size(); ++q) {
          SrcCell &m = ecells[q];
          if (m.comp == s.comp && m.k == s.k && m.j == s.j && m.i == s.i) {
            m.w += s.w;
            merged = true;
            break;
          }
        }
        if (!merged) ecells.push_back(s);
      }
      auto add_b = [&](int comp, int k, int j, int i) {
        if (!b_ok(comp, k, j, i)) return;
        for (size_t q = b0; q < bcells.size(); ++q) {
          const SrcCell &m = bcells[q];
          if (m.comp == comp && m.k == k && m.j == j && m.i == i) return;
        }
        bcells.push_back({comp, k, j, i, 0.0});
      };
      for (size_t q = e0; q < ecells.size(); ++q) {
        const SrcCell s = ecells[q];
        if (s.comp == 0) {  // Ex feeds By(k) and By(k-1)
          add_b(4, s.k, s.j, s.i);
          add_b(4, s.k - 1, s.j, s.i);
        } else if (s.comp == 1) {  // Ey feeds Bx(k) and Bx(k-1)
          add_b(3, s.k, s.j, s.i);
          add_b(3, s.k - 1, s.j, s.i);
        } else {  // Ez feeds Bx(j), Bx(j-1), By(i), By(i-1)
          add_b(3, s.k, s.j, s.i);
          add_b(3, s.k, s.j - 1, s.i);
          add_b(4, s.k, s.j, s.i);
          add_b(4, s.k, s.j, s.i - 1);
        }
      }
      std::sort(bcells.begin() + b0, bcells.end(), [](const SrcCell &x, const SrcCell &y) {
        if (x.comp != y.comp) return x.comp < y.comp;
        if (x.j != y.j) return x.j < y.j;
        if (x.i != y.i) return x.i < y.i;
        return x.k < y.k;
      });
      for (size_t q = b0; q < bcells.size(); ++q) {
        const SrcCell &x = bcells[q];
        const bool same = q > b0 && bcells[q - 1].comp == x.comp && bcells[q - 1].j == x.j && bcells[q - 1].i == x.i;
        if (same)
          ++runs.back().y;
        else
          runs.push_back(make_int2((int)q, 1));
      }
    }
    idx[nbatch + b] = (int)ecells.size() - idx[b];
    idx[3 * nbatch + b] = (int)runs.size() - idx[2 * nbatch + b];
  }
  c->has_fix = !ecells.empty();
  if (ecells.empty()) ecells.push_back({-1, -1, -1, -1, 0.0});
  if (bcells.empty()) bcells.push_back({-1, -1, -1, -1, 0.0});
  if (runs.empty()) runs.push_back(make_int2(0, 0));
  if (hipMalloc((void **)&c->d_fixe, ecells.size() * sizeof(SrcCell)) != hipSuccess ||
      hipMalloc((void **)&c->d_fixb, bcells.size() * sizeof(SrcCell)) != hipSuccess ||
      hipMalloc((void **)&c->d_fixg, runs.size() * sizeof(int2)) != hipSuccess ||
      hipMalloc((void **)&c->d_fixidx, idx.size() * sizeof(int)) != hipSuccess)
    return fail(TCSEM_ENOMEM, "source table allocation");
  HIPCHK(hipMemcpy(c->d_fixe, ecells.data(), ecells.size() * sizeof(SrcCell), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_fixb, bcells.data(), bcells.size() * sizeof(SrcCell), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_fixg, runs.data(), runs.size() * sizeof(int2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_fixidx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  return TCSEM_OK;
}

int tcsem_begin(tcsem_ctx *c, int tx_first, int nbatch) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  if (nbatch < 1) return fail(TCSEM_EINVAL, "batch must hold at least one transmitter");
  const bool have_tx = !c->tx.empty();
  if (have_tx && (tx_first < 0 || tx_first + nbatch > (int)c->tx.size()))
    return fail(TCSEM_EINVAL, "transmitters [%d, %d) outside the %zu set", tx_first, tx_first + nbatch, c->tx.size());
  int rc = tcsem_reserve_batch(c, nbatch);
  if (rc) return rc;
  free_run(c);
  c->tx_first = tx_first;
  c->nbatch = nbatch;
  choose_tiling(c, nbatch);
  if ((rc = tcsem_reset_fields(c, 0, nbatch))) return rc;
  // sources (per slot)
  std::vector<SrcCell> cells;
  std::vector<int> idx(2 * nbatch, 0);
  c->src_kmin = 1 << 30;
  c->src_kmax = -1;
  for (int b = 0; b < nbatch; ++b) {
    idx[b] = (int)cells.size();
    if (have_tx) {
      const TxLoc &t = c->tx[tx_first + b];
      for (const SrcCell &s : t.cells) {
        cells.push_back(s);
        c->src_kmin = std::min(c->src_kmin, s.k);
        c->src_kmax = std::max(c->src_kmax, s.k);
      }
    }
    idx[nbatch + b] = (int)cells.size() - idx[b];
  }
  if (cells.empty()) cells.push_back({-1, -1, -1, -1, 0.0});
  if (hipMalloc((void **)&c->d_src, cells.size() * sizeof(SrcCell)) != hipSuccess ||
      hipMalloc((void **)&c->d_srcidx, idx.size() * sizeof(int)) != hipSuccess)
    return fail(TCSEM_ENOMEM, "source table allocation");
  HIPCHK(hipMemcpy(c->d_src, cells.data(), cells.size() * sizeof(SrcCell), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_srcidx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  if ((rc = build_fix_lists(c, nbatch))) return rc;
  // receivers (per slot)
  c->nrmax = 1;
  if (c->have_rx) {
    std::vector<int> rl(2 * nbatch, 0);
    std::vector<int> loc;
    for (int b = 0; b < nbatch; ++b) {
      rl[b] = (int)loc.size();
      if (have_tx) {
        const TxLoc &t = c->tx[tx_first + b];
        loc.insert(loc.end(), t.rx.begin(), t.rx.end());
        rl[nbatch + b] = (int)t.rx.size();
        c->nrmax = std::max(c->nrmax, (int)t.rx.size());
      }
    }
    rl.insert(rl.end(), loc.begin(), loc.end());
    if (hipMalloc((void **)&c->d_rxl, std::max<size_t>(rl.size(), 1) * sizeof(int)) != hipSuccess)
      return fail(TCSEM_ENOMEM, "receiver list allocation");
    HIPCHK(hipMemcpy(c->d_rxl, rl.data(), rl.size() * sizeof(int), hipMemcpyHostToDevice));
    const size_t ne0 = (size_t)nbatch * 2 * c->ncomp * c->nrmax, nes = (size_t)nbatch * c->nrmax * c->ncomp * c->rt.Nt;
    if (hipMalloc((void **)&c->d_ef0, std::max<size_t>(ne0, 1) * 8) != hipSuccess ||
        hipMalloc((void **)&c->d_efs, std::max<size_t>(nes, 1) * 8) != hipSuccess ||
        hipMalloc((void **)&c->d_ops, (size_t)2 * c->rt.Nt * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&c->d_opl1, (size_t)2 * c->rt.Nt * sizeof(double)) != hipSuccess)
      return fail(TCSEM_ENOMEM, "trace allocation");
    // ef0 is zeroed (the reference leaves it uninitialised, main.cu:208)
    HIPCHK(launch_fill_zero(c->d_ef0, std::max<size_t>(ne0, 1) * 8, c->stream));
    HIPCHK(launch_fill_zero(c->d_efs, std::max<size_t>(nes, 1) * 8, c->stream));
  }
  // schedule state (main.cu:212-217), reset per batch
  const int ns = c->setup.n_src / 2;
  c->wt.assign(2 * ns + 1, 0);
  c->wdt.assign(2 * ns + 2, 0);
  c->wwave.assign(2 * ns + 1, 0);
  if ((rc = waveform(ns, c->setup.sigma_t, c->wt.data(), c->wdt.data(), c->wwave.data()))) return rc;
  c->n = 0;
  c->steps = 0;
  c->ne = c->nb = 0;
  std::memset(c->Tpof, 0, sizeof c->Tpof);
  c->toT = 0;
  c->dt1 = c->wdt[0];
  c->dt2 = c->wdt[1];
  c->Tpof[0][1] = c->toT + c->dt1;
  c->Tpof[1][1] = c->toT + c->dt1 + 0.5 * c->dt2;
  c->op_it.clear();
  c->op_l1.clear();
  c->up64 = upcont_args<double>(c);
  c->up32 = upcont_args<float>(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->begun = true;
  return TCSEM_OK;
}

int tcsem_step(tcsem_ctx *c, int nsteps) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  if (!c->begun) return fail(TCSEM_ESTATE, "tcsem_begin first");
  for (int q = 0; q < nsteps;) {
    const int r = c->prec == 64 ? advance<double>(c, nsteps - q) : advance<float>(c, nsteps - q);
    if (r < 0) return r;
    q += r;
  }
  return join_streams(c);
}

int tcsem_autotune(tcsem_ctx *c, int nbatch, int nsteps, int *rows, int *pf, int *zchunks, int *zsum, double *ms) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  if (nbatch < 1 || nsteps < 1 || nsteps > 1000) return fail(TCSEM_EINVAL, "batch >= 1 and 1 <= steps <= 1000");
  int rc = join_streams(c);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((rc = tcsem_reserve_batch(c, nbatch))) return rc;
  free_run(c);  // a begun run is discarded; tcsem_begin / tcsem_run start a new one
  if ((rc = tcsem_reset_fields(c, 0, nbatch))) return rc;
  int by = 0, p = 0, nch = 0, fuse = 1, s8 = 1;
  double t = 0.0;
  rc = c->prec == 64 ? autotune_march<double>(c, nbatch, nsteps, &by, &p, &nch, &fuse, &s8, &t)
                     : autotune_march<float>(c, nbatch, nsteps, &by, &p, &nch, &fuse, &s8, &t);
  if (rc) return rc;
  // two steps per pass, with the one-step tiling just picked for its top launches: kept when at
  // least 2% faster per step
  int pby = 0, ppf = 1;
  {
    const int fuse0 = c->fuse;
    const bool s80 = c->s8;
    apply_tiling(c, by, p, nch);
    c->fuse = fuse;
    c->s8 = s8 != 0;
    double pt = 0.0;
    rc = c->prec == 64 ? autotune_pairs<double>(c, nbatch, nsteps, &pby, &ppf, &pt)
                       : autotune_pairs<float>(c, nbatch, nsteps, &pby, &ppf, &pt);
    c->fuse = fuse0;
    c->s8 = s80;
    if (rc) return rc;
    if (pby && pt < 0.98 * t) {
      t = pt;
    } else {
      pby = 0;
      ppf = 1;
    }
  }
  c->tune_pair_by = pby;
  c->tune_pair_pf = ppf;
  if ((rc = tcsem_reset_fields(c, 0, nbatch))) return rc;
  c->tune_nbatch = nbatch;
  c->tune_by = by;
  c->tune_pf = p;
  c->tune_nch = nch;
  c->tune_fuse = fuse;
  c->tune_s8 = s8;
  choose_tiling(c, nbatch);
  if (rows) *rows = by;
  if (pf) *pf = p;
  if (zchunks) *zchunks = nch;
  if (zsum) *zsum = fuse;
  if (ms) *ms = t;
  return TCSEM_OK;
}

int tcsem_steps_remaining(const tcsem_ctx *c) {
  if (!c || !c->begun || !c->have_rx) return -1;
  return sampling_done(c) ? 0 : 1;
}

int tcsem_get_traces(tcsem_ctx *c, int slot, double *traces) {
  if (check_ctx(c) || !traces) return fail(TCSEM_EINVAL, "null argument");
  if (!c->begun || !c->have_rx) return fail(TCSEM_ESTATE, "no run in progress");
  if (slot < 0 || slot >= c->nbatch) return fail(TCSEM_EINVAL, "slot out of range");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int nr = c->tx.empty() ? 0 : (int)c->tx[c->tx_first + slot].rx.size();
  const size_t per = (size_t)c->nrmax * c->ncomp * c->rt.Nt;
  std::vector<double> h(per);
  HIPCHK(hipMemcpy(h.data(), c->d_efs + per * slot, per * 8, hipMemcpyDeviceToHost));
  std::memcpy(traces, h.data(), (size_t)nr * c->ncomp * c->rt.Nt * 8);
  return TCSEM_OK;
}

int tcsem_trace_shape(const tcsem_ctx *c, int slot, int *nr, int *ncomp, int *nt) {
  if (!c || !nr || !ncomp || !nt) return fail(TCSEM_EINVAL, "null argument");
  if (!c->begun || !c->have_rx) return fail(TCSEM_ESTATE, "no run in progress");
  if (slot < 0 || slot >= c->nbatch) return fail(TCSEM_EINVAL, "slot out of range");
  *nr = c->tx.empty() ? 0 : (int)c->tx[c->tx_first + slot].rx.size();
  *ncomp = c->ncomp;
  *nt = c->rt.Nt;
  return TCSEM_OK;
}

int tcsem_run(tcsem_ctx *c, int tx_first, int ntx, int max_steps, double *traces, int *steps_done) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  if (!c->have_rx || c->tx.empty()) return fail(TCSEM_ESTATE, "set receivers and transmitters first");
  int rc = tcsem_begin(c, tx_first, ntx);
  if (rc) return rc;
  while (!sampling_done(c) && (max_steps <= 0 || c->steps < max_steps)) {
    const long room = max_steps > 0 ? max_steps - c->steps : (long)1 << 40;
    const int r = c->prec == 64 ? advance<double>(c, room) : advance<float>(c, room);
    if (r < 0) return r;
  }
  if ((rc = join_streams(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (steps_done) *steps_done = c->steps;
  if (traces) {
    size_t off = 0;
    for (int b = 0; b < ntx; ++b) {
      if ((rc = tcsem_get_traces(c, b, traces + off))) return rc;
      off += (size_t)c->tx[tx_first + b].rx.size() * c->ncomp * c->rt.Nt;
    }
  }
  return TCSEM_OK;
}

int tcsem_get_field(tcsem_ctx *c, int slot, int comp, double *out) {
  if (check_ctx(c) || !out) return fail(TCSEM_EINVAL, "null argument");
  if (slot < 0 || slot >= c->nslots) return fail(TCSEM_EINVAL, "slot out of range");
  const int L = c->g.info.L, M = c->g.info.M, N = c->g.info.N;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (comp == TCSEM_BZ) {
    void *tmp = nullptr;
    const size_t n = (size_t)(N + 1) * M * L;
    if (hipMalloc(&tmp, n * c->esz) != hipSuccess) return fail(TCSEM_ENOMEM, "Bz read-back allocation");
    hipError_t e = c->prec == 64 ? launch_bz_volume<double>(march_args<double>(c), (double *)tmp, slot, c->stream)
                                 : launch_bz_volume<float>(march_args<float>(c), (float *)tmp, slot, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    std::vector<char> h(n * c->esz);
    if (e == hipSuccess) e = hipMemcpy(h.data(), tmp, n * c->esz, hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(TCSEM_EDEVICE, "Bz read-back: %s", hipGetErrorString(e));
    for (size_t q = 0; q < n; ++q) out[q] = c->prec == 64 ? ((double *)h.data())[q] : ((float *)h.data())[q];
    return TCSEM_OK;
  }
  int nk, nj, ni;
  const char *base;
  size_t bytes;
  if (comp >= TCSEM_EX && comp <= TCSEM_BY) {
    static const int dk[5] = {1, 1, 0, 0, 0}, dj[5] = {1, 0, 1, 0, 1}, di[5] = {0, 1, 1, 1, 0};
    nk = N + dk[comp];
    nj = M + dj[comp];
    ni = L + di[comp];
    base = c->d_fields + ((size_t)(c->cur * 5 + comp) * c->nslots + slot) * c->fstride * c->esz;
    bytes = (size_t)c->fstride * c->esz;
  } else if (comp == TCSEM_BXAIR || comp == TCSEM_BYAIR) {
    nk = 1;
    nj = comp == TCSEM_BXAIR ? M : M + 1;
    ni = comp == TCSEM_BXAIR ? L + 1 : L;
    base = c->d_planes + ((size_t)(comp == TCSEM_BXAIR ? 1 : 2) * c->nslots + slot) * c->pstride * c->esz;
    bytes = (size_t)c->pstride * c->esz;
  } else {
    return fail(TCSEM_EINVAL, "unknown component %d", comp);
  }
  std::vector<char> h(bytes);
  HIPCHK(hipMemcpy(h.data(), base, bytes, hipMemcpyDeviceToHost));
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        const size_t src = (size_t)k * c->PS + (size_t)(j + 1) * c->PX + (i + c->xm);
        out[((size_t)k * nj + j) * ni + i] = c->prec == 64 ? ((double *)h.data())[src] : ((float *)h.data())[src];
      }
  return TCSEM_OK;
}

int tcsem_set_field(tcsem_ctx *c, int slot, int comp, const double *in) {
  if (c) c->ib_valid = false;
  if (check_ctx(c) || !in) return fail(TCSEM_EINVAL, "null argument");
  if (slot < 0 || slot >= c->nslots) return fail(TCSEM_EINVAL, "slot out of range");
  const int L = c->g.info.L, M = c->g.info.M, N = c->g.info.N;
  int nk, nj, ni;
  char *base;
  size_t bytes;
  if (comp >= TCSEM_EX && comp <= TCSEM_BY) {
    static const int dk[5] = {1, 1, 0, 0, 0}, dj[5] = {1, 0, 1, 0, 1}, di[5] = {0, 1, 1, 1, 0};
    nk = N + dk[comp];
    nj = M + dj[comp];
    ni = L + di[comp];
    base = c->d_fields + ((size_t)(c->cur * 5 + comp) * c->nslots + slot) * c->fstride * c->esz;
    bytes = (size_t)c->fstride * c->esz;
  } else if (comp == TCSEM_BXAIR || comp == TCSEM_BYAIR) {
    nk = 1;
    nj = comp == TCSEM_BXAIR ? M : M + 1;
    ni = comp == TCSEM_BXAIR ? L + 1 : L;
    base = c->d_planes + ((size_t)(comp == TCSEM_BXAIR ? 1 : 2) * c->nslots + slot) * c->pstride * c->esz;
    bytes = (size_t)c->pstride * c->esz;
  } else {
    return fail(TCSEM_EINVAL, "component %d is not settable (Bz is derived from Bx, By)", comp);
  }
  std::vector<char> h(bytes, 0);
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        const size_t dst = (size_t)k * c->PS + (size_t)(j + 1) * c->PX + (i + c->xm);
        const double v = in[((size_t)k * nj + j) * ni + i];
        if (c->prec == 64) ((double *)h.data())[dst] = v;
        else ((float *)h.data())[dst] = (float)v;
      }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(base, h.data(), bytes, hipMemcpyHostToDevice));
  return TCSEM_OK;
}

int tcsem_upcont_plane(tcsem_ctx *c, const double *bz0, double *bxair, double *byair) {
  if (check_ctx(c) || !bz0 || !bxair || !byair) return fail(TCSEM_EINVAL, "null argument");
  const int L = c->g.info.L, M = c->g.info.M;
  std::vector<char> h((size_t)c->pstride * c->esz, 0);
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < L; ++i) {
      const size_t dst = (size_t)(j + 1) * c->PX + (i + c->xm);
      if (c->prec == 64) ((double *)h.data())[dst] = bz0[(size_t)j * L + i];
      else ((float *)h.data())[dst] = (float)bz0[(size_t)j * L + i];
    }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->d_planes, h.data(), h.size(), hipMemcpyHostToDevice));
  hipError_t e = c->prec == 64 ? launch_upcont<double>(c->up64, 1, c->stream) : launch_upcont<float>(c->up32, 1, c->stream);
  if (e != hipSuccess) return fail(TCSEM_EDEVICE, "upward continuation: %s", hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = tcsem_get_field(c, 0, TCSEM_BXAIR, bxair);
  if (rc) return rc;
  return tcsem_get_field(c, 0, TCSEM_BYAIR, byair);
}

extern "C++" {
namespace {
// The tiled chain of tcsem_upcont_plane on caller device planes (dense rows), with its work planes
// (one slot's A, G, Bx and By regions, the layout of the context's own work buffer) in the
// library's scratch array at scr.  Every work entry the chain reads is written earlier in the same
// call, so nothing is cleared.
template <typename T>
hipError_t upcont_on_planes(tcsem_ctx *c, void *scr, const void *bz0, void *bxair, void *byair) {
  UpcontArgs<T> u = up_of<T>(c);
  const int L = c->g.info.L;
  u.bz0 = (const T *)bz0;
  u.bxair = (T *)bxair;
  u.byair = (T *)byair;
  u.pstride = 0;
  u.pz = L;
  u.pxa = L + 1;
  u.pya = L;
  T *w = (T *)scr;
  u.A = w;
  u.G = w + 2 * c->sA;
  u.tbx = w + 2 * (c->sA + c->sG);
  u.tby = u.tbx + c->sX;
  hipError_t e = launch_upcont<T>(u, 1, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}
}  // namespace
}  // extern "C++"

int tcsem_upcont_plane_noalloc(tcsem_ctx *c, const void *d_bz0, void *d_bxair, void *d_byair) {
  if (check_ctx(c) || !d_bz0 || !d_bxair || !d_byair) return fail(TCSEM_EINVAL, "null argument");
  if (!c->up_tiled)
    return fail(TCSEM_ESIZE, "air plane %d x %d: FFT lengths above 2048 have no allocation-free path", c->g.info.L1,
                c->g.info.M1);
  const size_t need = (size_t)(2 * c->sA + 2 * c->sG + 2 * c->sX) * c->esz;
  if (need > kUpcontScratchBytes)
    return fail(TCSEM_ESIZE, "air plane %d x %d needs %zu bytes of work planes; the allocation-free limit is %zu",
                c->g.info.L1, c->g.info.M1, need, kUpcontScratchBytes);
  void *scr = upcont_scratch(c->dev);
  if (!scr) return fail(TCSEM_EDEVICE, "no scratch address for device %d", c->dev);
  const hipError_t e = c->prec == 64 ? upcont_on_planes<double>(c, scr, d_bz0, d_bxair, d_byair)
                                     : upcont_on_planes<float>(c, scr, d_bz0, d_bxair, d_byair);
  if (e != hipSuccess) return fail(TCSEM_EDEVICE, "upward continuation: %s", hipGetErrorString(e));
  return TCSEM_OK;
}

int tcsem_enable_timing(tcsem_ctx *c, int on) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  if (on) {
    int rc = rc_collect(c);
    if (rc) return rc;
    for (int q = 0; q < 5; ++q) {
      c->tacc[q] = 0;
      c->tn[q] = 0;
    }
  } else {
    int rc = rc_collect(c);  // keep what was recorded readable
    if (rc) return rc;
  }
  c->timing = on != 0;
  return TCSEM_OK;
}

int tcsem_timing(const tcsem_ctx *cc, double t[4], long n[4]) {
  if (!cc || !t || !n) return fail(TCSEM_EINVAL, "null argument");
  tcsem_ctx *c = const_cast<tcsem_ctx *>(cc);  // folding recorded intervals does not change the state
  int rc = rc_collect(c);
  if (rc) return rc;
  for (int q = 0; q < 4; ++q) {
    n[q] = c->tn[q];
    t[q] = c->tn[q] ? c->tacc[q] / c->tn[q] : 0.0;
  }
  return TCSEM_OK;
}

int tcsem_pair_timing(const tcsem_ctx *cc, double *ms, long *n) {
  if (!cc || !ms || !n) return fail(TCSEM_EINVAL, "null argument");
  tcsem_ctx *c = const_cast<tcsem_ctx *>(cc);
  int rc = rc_collect(c);
  if (rc) return rc;
  *n = c->tn[4];
  *ms = c->tn[4] ? c->tacc[4] / c->tn[4] : 0.0;
  return TCSEM_OK;
}

int tcsem_set_pairs(tcsem_ctx *c, int rows, int pf) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  if (rows != 0 && (!pair_rows_supported(c->prec, rows) || pf < 1 || pf > 2))
    return fail(TCSEM_EINVAL, "two-step pass: rows 0 (off), 8, 10, 12, 14 or 16 (FP64: 8, 10, 11 or 12); prefetch 1 or 2");
  if (rows && c->g.info.N < 6) return fail(TCSEM_EINVAL, "two-step pass needs at least 6 levels");
  c->pair_set = rows;
  c->pair_set_pf = pf;
  pair_tiling(c, rows, pf);
  return TCSEM_OK;
}

int tcsem_pair_info(const tcsem_ctx *c, int *rows, int *pf, long *pairs) {
  if (check_ctx(c) || !rows || !pf || !pairs) return fail(TCSEM_EINVAL, "null argument");
  *rows = c->pair_by;
  *pf = c->pair_pf;
  *pairs = c->pairs_done;
  return TCSEM_OK;
}

int tcsem_sync(tcsem_ctx *c) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  HIPCHK(hipStreamSynchronize(c->stream));
  return TCSEM_OK;
}

// ---- model batch -------------------------------------------------------------------------------
int tcsem_batch_create(tcsem_ctx *const *models, int nmodel, tcsem_batch **out) {
  if (!models || !out || nmodel < 1) return fail(TCSEM_EINVAL, "null argument or empty batch");
  *out = nullptr;
  for (int q = 0; q < nmodel; ++q) {
    const tcsem_ctx *c = models[q];
    if (!c) return fail(TCSEM_EINVAL, "null member %d", q);
    if (!c->begun || c->nbatch != 1)
      return fail(TCSEM_ESTATE, "member %d: begin it with one transmitter (tcsem_begin(ctx, tx, 1)) first", q);
    if (c->prec != models[0]->prec) return fail(TCSEM_EINVAL, "members differ in precision");
    if (c->g.info.N < 6) return fail(TCSEM_ESIZE, "member %d: a batch needs at least 6 levels", q);
    if (!c->up_tiled) return fail(TCSEM_ESIZE, "member %d: air plane beyond the tiled chain (FFT length > 2048)", q);
    for (int r = 0; r < q; ++r)
      if (models[r] == c) return fail(TCSEM_EINVAL, "member %d appears twice", q);
  }
  tcsem_batch *b = new (std::nothrow) tcsem_batch();
  if (!b) return fail(TCSEM_ENOMEM, "batch allocation");
  b->m.assign(models, models + nmodel);
  b->prec = models[0]->prec;
  b->active.assign(nmodel, 1);
  const size_t rsz = b->prec == 64 ? sizeof(MarchArgs<double>) : sizeof(MarchArgs<float>);
  const size_t usz = b->prec == 64 ? sizeof(UpcontArgs<double>) : sizeof(UpcontArgs<float>);
  const size_t psz = b->prec == 64 ? sizeof(PairArgs<double>) : sizeof(PairArgs<float>);
  bool ok = hipStreamCreate(&b->stream) == hipSuccess && hipStreamCreate(&b->stream2) == hipSuccess &&
            hipEventCreateWithFlags(&b->xev[0], hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->xev[1], hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess &&
            hipMalloc((void **)&b->d_geo, 2 * nmodel * rsz) == hipSuccess &&
            hipMalloc((void **)&b->d_ugeo, nmodel * usz) == hipSuccess &&
            hipMalloc((void **)&b->d_pgeo, 2 * nmodel * psz) == hipSuccess &&
            hipMalloc((void **)&b->d_tgeo, 4 * nmodel * rsz) == hipSuccess;
  if (!ok) {
    tcsem_batch_destroy(b);
    return fail(TCSEM_EDEVICE, "batch stream, event or record allocation failed");
  }
  *out = b;
  return TCSEM_OK;
}

void tcsem_batch_destroy(tcsem_batch *b) {
  if (!b) return;
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  (void)hipFree(b->d_geo);
  (void)hipFree(b->d_ugeo);
  (void)hipFree(b->d_scal);
  (void)hipFree(b->d_work);
  (void)hipFree(b->d_pgeo);
  (void)hipFree(b->d_tgeo);
  (void)hipFree(b->d_pwork);
  for (auto &e : b->xev)
    if (e) (void)hipEventDestroy(e);
  if (b->done) (void)hipEventDestroy(b->done);
  for (auto &e : b->evpool)
    if (e) (void)hipEventDestroy(e);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int tcsem_batch_autotune(tcsem_batch *b, int nsteps, int *rows, int *pf, double *ms) {
  if (!b) return fail(TCSEM_EINVAL, "null batch");
  if (nsteps < 1 || nsteps > 1000) return fail(TCSEM_EINVAL, "1 <= steps <= 1000");
  for (const tcsem_ctx *c : b->m)
    if (int rc = batch_fail_if(c->steps != 0, "autotune a batch before its first step")) return rc;
  int by = 0, p = 0, s8 = 0, pr = 0;
  double t = 0.0;
  const int rc = b->prec == 64 ? batch_autotune_t<double>(b, nsteps, &by, &p, &s8, &pr, &t)
                               : batch_autotune_t<float>(b, nsteps, &by, &p, &s8, &pr, &t);
  if (rc) return rc;
  b->by = by;
  b->pf = p;
  b->s8 = s8 != 0;
  b->pair_tuned = pr;
  if (rows) *rows = by;
  if (pf) *pf = p;
  if (ms) *ms = t;
  return TCSEM_OK;
}

int tcsem_batch_step(tcsem_batch *b, int nsteps) {
  if (!b) return fail(TCSEM_EINVAL, "null batch");
  if (nsteps < 0) return fail(TCSEM_EINVAL, "negative step count");
  if (nsteps == 0) return TCSEM_OK;
  for (const tcsem_ctx *c : b->m)
    if (int rc = batch_fail_if(!c->begun || c->nbatch != 1, "a member's run was restarted or ended")) return rc;
  return b->prec == 64 ? batch_step_t<double>(b, nsteps) : batch_step_t<float>(b, nsteps);
}

int tcsem_batch_run(tcsem_batch *b, int max_steps, int *steps_done) {
  if (!b) return fail(TCSEM_EINVAL, "null batch");
  constexpr int kChunk = 256;  // steps per call of the step loop (members finish at a chunk's end)
  for (;;) {
    bool any = false;
    int most = 0;
    for (const tcsem_ctx *c : b->m) {
      if (!c->have_rx) return fail(TCSEM_ESTATE, "set receivers of every member first");
      if (!sampling_done(c)) any = true;
      most = std::max(most, c->steps);
    }
    if (!any || (max_steps > 0 && most >= max_steps)) break;
    const int n = max_steps > 0 ? std::min(kChunk, max_steps - most) : kChunk;
    if (int rc = tcsem_batch_step(b, n)) return rc;
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  if (steps_done) *steps_done = b->steps;
  return TCSEM_OK;
}

int tcsem_set_graph(tcsem_ctx *c, int nsteps) {
  if (check_ctx(c)) return TCSEM_EINVAL;
  if (nsteps != 0 && (nsteps < 2 || nsteps > kGraphMaxSteps || (nsteps & 1)))
    return fail(TCSEM_EINVAL, "graph steps: 0 (off) or even, 2 .. %d", kGraphMaxSteps);
  if (nsteps && !c->d_gs &&
      hipMalloc((void **)&c->d_gs, (size_t)4 * kGraphMaxSteps * sizeof(double)) != hipSuccess)
    return fail(TCSEM_ENOMEM, "graph scalar allocation");
  if (int rc = join_streams(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->graph_k = nsteps;
  return TCSEM_OK;
}

long tcsem_graph_replays(const tcsem_ctx *c) { return c ? c->graph_replays : -1; }

int tcsem_sigma_table(const tcsem_ctx *c) { return c ? (c->s8 ? c->sig8_n : 0) : -1; }

int tcsem_batch_enable_timing(tcsem_batch *b, int on) {
  if (!b) return fail(TCSEM_EINVAL, "null batch");
  if (int rc = batch_collect(b)) return rc;
  b->timing = on != 0;
  for (int k = 0; k < 4; ++k) {
    b->tacc[k] = 0;
    b->tn[k] = 0;
  }
  return TCSEM_OK;
}

int tcsem_batch_set_pairs(tcsem_batch *b, int rows) {
  if (!b) return fail(TCSEM_EINVAL, "null batch");
  if (rows != 0 && !pair_rows_supported(b->prec, rows))
    return fail(TCSEM_EINVAL, "two-step pass: rows 0 (off), 8, 10, 12, 14 or 16 (FP64: 8, 10, 11 or 12)");
  b->pair_set = rows;
  return TCSEM_OK;
}

int tcsem_batch_pair_info(const tcsem_batch *b, int *rows, int *tuned, long *pairs) {
  if (!b || !rows || !tuned || !pairs) return fail(TCSEM_EINVAL, "null argument");
  *rows = b->pair_by;
  *tuned = b->pair_tuned;
  *pairs = b->pairs_done;
  return TCSEM_OK;
}

int tcsem_batch_pair_timing(tcsem_batch *b, double *ms, long *n) {
  if (!b || !ms || !n) return fail(TCSEM_EINVAL, "null argument");
  if (int rc = batch_collect(b)) return rc;
  *n = b->tn[3];
  *ms = b->tn[3] ? b->tacc[3] / b->tn[3] : 0.0;
  return TCSEM_OK;
}

int tcsem_batch_timing(tcsem_batch *b, double t[3], long n[3]) {
  if (!b || !t || !n) return fail(TCSEM_EINVAL, "null argument");
  if (int rc = batch_collect(b)) return rc;
  for (int k = 0; k < 3; ++k) {
    t[k] = b->tn[k] ? b->tacc[k] / b->tn[k] : 0.0;
    n[k] = b->tn[k];
  }
  return TCSEM_OK;
}

int tcsem_batch_sync(tcsem_batch *b) {
  if (!b) return fail(TCSEM_EINVAL, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream2));
  HIPCHK(hipStreamSynchronize(b->stream));
  return TCSEM_OK;
}

}  // extern "C"
