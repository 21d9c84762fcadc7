// tcsem_kernels.hip -- gfx950 kernels of the DuFort-Frankel step: the fused E/B volume march,
// the surface level, the derived Bz volume and the receiver gather.
//
// Volume march (levels N-1 .. 1), one workgroup per 62 x 6 column tile, one lane per (i, j)
// column, marching z bottom-up.  Per level it
//   (a) extends the running Bz integral of the *old* B (Bderivate.cu:227-253): Bz is never
//       stored, the E update reads it from this register sum and its LDS copy;
//   (b) prefetches B, sigma and E of the next level up into registers / the other LDS buffer;
//   (c) updates Ez (ElectricalFiled.cu:277-345) on the tile plus a +x/+y halo;
//   (d) updates Ex, Ey (ElectricalFiled.cu:89-274) and applies the bipole cells
//       (ElectricalFiled.cu:347-475);
//   (e) updates Bx, By (Bderivate.cu:79-200) from the new E of this level and the level below.
// A level reads B_old at k and k-1 and E_new at k and k+1, so one bottom-up pass moves every
// field once: read sigma, E(3), Bx, By and write E(3), Bx, By -- 11 words per cell against the
// 13 of a stored-Bz design.  The halo ring (one column/row each side) keeps every neighbour
// read inside the workgroup's LDS; halo lanes recompute Bz sums (-x/-y side) or Ez (+x/+y).
// Level 0 needs the air fields from the upward continuation of Bz(0), so the march stops at
// level 1 and writes Bz_old(0); k_surface finishes level 0 after the air planes are ready.
#include "tcsem_internal.h"

namespace tcsem {

// Running Bz integral: plain FP64, or compensated (TwoSum) FP32.
template <typename T>
struct Acc;
template <>
struct Acc<double> {
  double s;
  __device__ void init() { s = 0.0; }
  __device__ void add(double x) { s += x; }
  __device__ double val() const { return s; }
};
template <>
struct Acc<float> {
  float s, c;
  __device__ void init() { s = 0.f; c = 0.f; }
  __device__ void add(float x) {
    const float t = s + x;
    const float bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  __device__ float val() const { return s + c; }
};

// DuFort-Frankel update E <- (2l - s dt)/(2l + s dt) E + 4 dt/(2l + s dt) curl / mu0
// (ElectricalFiled.cu:138-139).  FP64 fields: the reference expression.  FP32 fields: the two
// coefficients in FP64, the product in FP32.
__device__ inline double df_new(double e, double sig, double lam, double dt, double c4, double curl) {
  (void)c4;
  const double den = 2.0 * lam + sig * dt;
  return (2.0 * lam - sig * dt) / den * e + 4.0 * dt / den * curl / kMu0;
}
__device__ inline float df_new(float e, float sig, double lam, double dt, double c4, float curl) {
  const double sd = (double)sig * dt;
  const double r = 1.0 / (2.0 * lam + sd);
  const float ca = (float)((2.0 * lam - sd) * r);
  const float cb = (float)(c4 * r);
  return ca * e + cb * curl;
}

template <typename T>
__device__ inline T apply_src(T e, int comp, int k, int j, int i, const MarchArgs<T> &a,
                              int first, int cnt) {
  if (k < a.src_kmin || k > a.src_kmax || a.amp == 0.0) return e;
  double w = 0.0;
  for (int c = 0; c < cnt; ++c) {
    const SrcCell s = a.src[first + c];
    if (s.comp == comp && s.k == k && s.j == j && s.i == i) w += s.w;
  }
  return (w != 0.0) ? (T)((double)e - a.amp * w) : e;
}

struct Tile {
  int lx, ly, i, j;
  bool inside, own, need_bz, need_ez;
};

template <int BY>
__device__ inline bool tile_of(int ntx, int nty, int nper, Tile &t) {
  constexpr int TX = kMarchLanes - 2, TY = BY - 2;
  // Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous band of tiles so
  // the halo re-reads of neighbouring tiles hit the same L2.
  const int b = blockIdx.x;
  const int tile = (b & 7) * nper + (b >> 3);
  if (tile >= ntx * nty) return false;
  const int tx = tile % ntx, ty = tile / ntx;
  t.lx = threadIdx.x;
  t.ly = threadIdx.y;
  t.i = tx * TX - 1 + t.lx;
  t.j = ty * TY - 1 + t.ly;
  const bool inx = t.lx >= 1 && t.lx <= TX, iny = t.ly >= 1 && t.ly <= TY;
  t.own = inx && iny;
  const bool bzh = (t.lx == 0 && iny) || (t.ly == 0 && inx);
  const bool ezh = (t.lx == kMarchLanes - 1 && iny) || (t.ly == BY - 1 && inx);
  t.need_bz = t.own || bzh;
  t.need_ez = t.own || ezh;
  return true;
}

template <typename T, int BY>
__global__ void __launch_bounds__(kMarchLanes *BY) k_march(const MarchArgs<T> a) {
  __shared__ T sBx[2][BY][kMarchLanes], sBy[2][BY][kMarchLanes], sS[2][BY][kMarchLanes];
  __shared__ T sBz[BY][kMarchLanes], sEz[BY][kMarchLanes];
  Tile t;
  if (!tile_of<BY>(a.ntx, a.nty, a.nper, t)) return;
  const int lx = t.lx, ly = t.ly, i = t.i, j = t.j;
  const int L = a.L, M = a.M, N = a.N;
  const int slot = blockIdx.y;
  const long long fo = (long long)slot * a.fstride, PS = a.PS;
  const T *__restrict__ exi = a.exi + fo;
  const T *__restrict__ eyi = a.eyi + fo;
  const T *__restrict__ ezi = a.ezi + fo;
  const T *__restrict__ bxi = a.bxi + fo;
  const T *__restrict__ byi = a.byi + fo;
  T *__restrict__ ex = a.ex + fo;
  T *__restrict__ ey = a.ey + fo;
  T *__restrict__ ez = a.ez + fo;
  T *__restrict__ bx = a.bx + fo;
  T *__restrict__ by = a.by + fo;
  const T *__restrict__ sig = a.sig;
  const bool inside = i >= -1 && i <= L + 1 && j >= -1 && j <= M + 1;
  const bool own = t.own && i <= L && j <= M;
  const bool vEx = own && i < L && j >= 1 && j < M;
  const bool vEy = own && i >= 1 && i < L && j < M;
  const bool vEzAny = t.need_ez && i >= 1 && i < L && j >= 1 && j < M;
  const bool vEz = own && vEzAny;
  const bool vBx = own && i >= 1 && i < L && j < M;
  const bool vBy = own && i < L && j >= 1 && j < M;
  const long long col = (long long)j * a.PX + i;
  T rdx = 0, rdx2 = 0, rdy = 0, rdy2 = 0;
  if (inside) {
    rdx = a.rdx[i + 1];
    rdx2 = a.rdx2[i + 1];
    rdy = a.rdy[j + 1];
    rdy2 = a.rdy2[j + 1];
  }
  const int sfirst = a.src_first ? a.src_first[slot] : 0, scount = a.src_count ? a.src_count[slot] : 0;
  const double lam = a.lam, dt = a.dt, c4 = 4.0 * a.dt / kMu0;
  const T dtb = (T)a.dtb;

  // level N-1 state
  T bx_k = 0, by_k = 0, s_k = 0, exo = 0, eyo = 0, ezo = 0;
  {
    const long long o = (long long)(N - 1) * PS + col;
    if (inside) {
      bx_k = bxi[o];
      by_k = byi[o];
      s_k = sig[o];
    }
    if (vEx) exo = exi[o];
    if (vEy) eyo = eyi[o];
    if (vEzAny) ezo = ezi[o];
  }
  int p = (N - 1) & 1;
  sBx[p][ly][lx] = bx_k;
  sBy[p][ly][lx] = by_k;
  sS[p][ly][lx] = s_k;
  Acc<T> bz;
  bz.init();
  T ex_n1 = 0, ey_n1 = 0;  // E_new at level k+1 (zero at the bottom wall, level N)
  __syncthreads();

  for (int k = N - 1; k >= 1; --k) {
    p = k & 1;
    const int q = p ^ 1;
    const T dzk = a.dzt[k], rdzk = a.rdz[k], rdz2k = a.rdz2[k];
    // (a) Bz_old(k) = Bz_old(k+1) + dz(k) div_h(B_old(k))
    if (t.need_bz) bz.add(dzk * ((sBx[p][ly][lx + 1] - bx_k) * rdx + (sBy[p][ly + 1][lx] - by_k) * rdy));
    const T bzk = bz.val();
    sBz[ly][lx] = bzk;
    // (b) next level up
    T bx_m = 0, by_m = 0, s_m = 0, exo_n = 0, eyo_n = 0, ezo_n = 0;
    {
      const long long o = (long long)(k - 1) * PS + col;
      if (inside) {
        bx_m = bxi[o];
        by_m = byi[o];
        s_m = sig[o];
      }
      if (vEx) exo_n = exi[o];
      if (vEy) eyo_n = eyi[o];
      if (vEzAny) ezo_n = ezi[o];
    }
    sBx[q][ly][lx] = bx_m;
    sBy[q][ly][lx] = by_m;
    sS[q][ly][lx] = s_m;
    __syncthreads();
    const long long ok = (long long)k * PS + col;
    // (c) Ez(k)
    T ezn = 0;
    if (vEzAny) {
      const T s4 = (T)0.25 * (sS[p][ly - 1][lx - 1] + sS[p][ly - 1][lx] + sS[p][ly][lx - 1] + sS[p][ly][lx]);
      const T curl = (by_k - sBy[p][ly][lx - 1]) * rdx2 - (bx_k - sBx[p][ly - 1][lx]) * rdy2;
      ezn = df_new(ezo, s4, lam, dt, c4, curl);
      ezn = apply_src(ezn, 2, k, j, i, a, sfirst, scount);
      if (vEz) ez[ok] = ezn;
    }
    sEz[ly][lx] = ezn;
    // (d) Ex(k), Ey(k)
    T exn = 0, eyn = 0;
    if (vEx) {
      const T sxk = (T)0.5 * (sS[p][ly - 1][lx] + sS[p][ly][lx]);
      const T sxm = (T)0.5 * (sS[q][ly - 1][lx] + sS[q][ly][lx]);
      const T ave = (T)0.5 * (sxk + sxm);
      const T curl = (bzk - sBz[ly - 1][lx]) * rdy2 - (by_k - by_m) * rdz2k;
      exn = df_new(exo, ave, lam, dt, c4, curl);
      exn = apply_src(exn, 0, k, j, i, a, sfirst, scount);
      ex[ok] = exn;
    }
    if (vEy) {
      const T s4 = (T)0.25 * (sS[p][ly][lx - 1] + sS[p][ly][lx] + sS[q][ly][lx - 1] + sS[q][ly][lx]);
      const T curl = (bx_k - bx_m) * rdz2k - (bzk - sBz[ly][lx - 1]) * rdx2;
      eyn = df_new(eyo, s4, lam, dt, c4, curl);
      eyn = apply_src(eyn, 1, k, j, i, a, sfirst, scount);
      ey[ok] = eyn;
    }
    __syncthreads();
    // (e) Bx(k), By(k) from E_new(k), E_new(k+1)
    if (vBx) bx[ok] = bx_k + dtb * ((ey_n1 - eyn) * rdzk - (sEz[ly + 1][lx] - ezn) * rdy);
    if (vBy) by[ok] = by_k + dtb * ((sEz[ly][lx + 1] - ezn) * rdx - (ex_n1 - exn) * rdzk);
    ex_n1 = exn;
    ey_n1 = eyn;
    bx_k = bx_m;
    by_k = by_m;
    s_k = s_m;
    exo = exo_n;
    eyo = eyo_n;
    ezo = ezo_n;
  }
  // Bz_old(0) for the upward continuation and the surface level
  if (t.need_bz) bz.add(a.dzt[0] * ((sBx[0][ly][lx + 1] - bx_k) * rdx + (sBy[0][ly + 1][lx] - by_k) * rdy));
  if (own && i >= 0 && i < L && j >= 0 && j < M) a.bz0[(long long)slot * a.pstride + col] = bz.val();
}

// Level 0: the air cells count as sigma = 0 and the level above is BxAir/ByAir at -dz0/2
// (ElectricalFiled.cu:107-139, 204-234).
template <typename T, int BY>
__global__ void __launch_bounds__(kMarchLanes *BY) k_surface(const MarchArgs<T> a) {
  __shared__ T sBx[BY][kMarchLanes], sBy[BY][kMarchLanes], sS[BY][kMarchLanes];
  __shared__ T sBz[BY][kMarchLanes], sEz[BY][kMarchLanes];
  Tile t;
  if (!tile_of<BY>(a.ntx, a.nty, a.nper, t)) return;
  const int lx = t.lx, ly = t.ly, i = t.i, j = t.j;
  const int L = a.L, M = a.M;
  const int slot = blockIdx.y;
  const long long fo = (long long)slot * a.fstride, po = (long long)slot * a.pstride, PS = a.PS;
  const T *__restrict__ exi = a.exi + fo;
  const T *__restrict__ eyi = a.eyi + fo;
  const T *__restrict__ ezi = a.ezi + fo;
  const T *__restrict__ bxi = a.bxi + fo;
  const T *__restrict__ byi = a.byi + fo;
  T *__restrict__ ex = a.ex + fo;
  T *__restrict__ ey = a.ey + fo;
  T *__restrict__ ez = a.ez + fo;
  T *__restrict__ bx = a.bx + fo;
  T *__restrict__ by = a.by + fo;
  const bool inside = i >= -1 && i <= L + 1 && j >= -1 && j <= M + 1;
  const bool own = t.own && i <= L && j <= M;
  const bool vEx = own && i < L && j >= 1 && j < M;
  const bool vEy = own && i >= 1 && i < L && j < M;
  const bool vEzAny = t.need_ez && i >= 1 && i < L && j >= 1 && j < M;
  const bool vEz = own && vEzAny;
  const bool vBx = own && i >= 1 && i < L && j < M;
  const bool vBy = own && i < L && j >= 1 && j < M;
  const long long col = (long long)j * a.PX + i;
  T rdx = 0, rdx2 = 0, rdy = 0, rdy2 = 0, bx0 = 0, by0 = 0, s0 = 0, bzc = 0;
  if (inside) {
    rdx = a.rdx[i + 1];
    rdx2 = a.rdx2[i + 1];
    rdy = a.rdy[j + 1];
    rdy2 = a.rdy2[j + 1];
    bx0 = bxi[col];
    by0 = byi[col];
    s0 = a.sig[col];
  }
  if (t.need_bz && i >= 0 && i < L && j >= 0 && j < M) bzc = a.bz0[po + col];
  T exo = 0, eyo = 0, ezo = 0, ex1 = 0, ey1 = 0, bxa = 0, bya = 0;
  if (vEx) {
    exo = exi[col];
    ex1 = ex[PS + col];  // E_new(1), written by the march
    bya = a.byair[po + col];
  }
  if (vEy) {
    eyo = eyi[col];
    ey1 = ey[PS + col];
    bxa = a.bxair[po + col];
  }
  if (vEzAny) ezo = ezi[col];
  sBx[ly][lx] = bx0;
  sBy[ly][lx] = by0;
  sS[ly][lx] = s0;
  sBz[ly][lx] = bzc;
  const int sfirst = a.src_first ? a.src_first[slot] : 0, scount = a.src_count ? a.src_count[slot] : 0;
  const double lam = a.lam, dt = a.dt, c4 = 4.0 * a.dt / kMu0;
  const T dtb = (T)a.dtb, rdz0 = a.rdz[0], rdz20 = a.rdz2[0];
  const T zero = 0;
  __syncthreads();
  T ezn = 0;
  if (vEzAny) {
    const T s4 = (T)0.25 * (sS[ly - 1][lx - 1] + sS[ly - 1][lx] + sS[ly][lx - 1] + sS[ly][lx]);
    const T curl = (by0 - sBy[ly][lx - 1]) * rdx2 - (bx0 - sBx[ly - 1][lx]) * rdy2;
    ezn = df_new(ezo, s4, lam, dt, c4, curl);
    ezn = apply_src(ezn, 2, 0, j, i, a, sfirst, scount);
    if (vEz) ez[col] = ezn;
  }
  sEz[ly][lx] = ezn;
  T exn = 0, eyn = 0;
  if (vEx) {
    const T sxk = (T)0.5 * (sS[ly - 1][lx] + sS[ly][lx]);
    const T ave = (T)0.5 * (sxk + zero);
    const T curl = (bzc - sBz[ly - 1][lx]) * rdy2 - (by0 - bya) * rdz20;
    exn = df_new(exo, ave, lam, dt, c4, curl);
    exn = apply_src(exn, 0, 0, j, i, a, sfirst, scount);
    ex[col] = exn;
  }
  if (vEy) {
    const T s4 = (T)0.25 * (sS[ly][lx - 1] + sS[ly][lx]);
    const T curl = (bx0 - bxa) * rdz20 - (bzc - sBz[ly][lx - 1]) * rdx2;
    eyn = df_new(eyo, s4, lam, dt, c4, curl);
    eyn = apply_src(eyn, 1, 0, j, i, a, sfirst, scount);
    ey[col] = eyn;
  }
  __syncthreads();
  if (vBx) bx[col] = bx0 + dtb * ((ey1 - eyn) * rdz0 - (sEz[ly + 1][lx] - ezn) * rdy);
  if (vBy) by[col] = by0 + dtb * ((sEz[ly][lx + 1] - ezn) * rdx - (ex1 - exn) * rdz0);
}

// Derived Bz volume (N+1, M, L dense) of one slot (Bderivate.cu:202-265), for field read-back.
template <typename T>
__global__ void k_bz_volume(const MarchArgs<T> a, T *bzout, int slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int L = a.L, M = a.M, N = a.N;
  if (i >= L || j >= M) return;
  const long long fo = (long long)slot * a.fstride;
  const T *bx = a.bxi + fo, *by = a.byi + fo;
  const T rdx = a.rdx[i + 1], rdy = a.rdy[j + 1];
  Acc<T> acc;
  acc.init();
  bzout[((long long)N * M + j) * L + i] = 0;
  for (int k = N - 1; k >= 0; --k) {
    const long long o = (long long)k * a.PS + (long long)j * a.PX + i;
    acc.add(a.dzt[k] * ((bx[o + 1] - bx[o]) * rdx + (by[o + a.PX] - by[o]) * rdy));
    bzout[((long long)k * M + j) * L + i] = acc.val();
  }
}

// ---- receivers (EFInterp2Rx.cu:46-470) ---------------------------------------------------------
// Natural cubic spline through 4 points evaluated on the middle interval (CubicSpline_1d4p).
__device__ inline double spline4(const double *x, const double *y, double t) {
  const double h1 = x[2] - x[1];
  const double v1 = 2.0 * (x[2] - x[0]);
  const double v2 = 2.0 * (x[3] - x[1]);
  const double u1 = 6.0 * ((y[2] - y[1]) / h1 - (y[1] - y[0]) / (x[1] - x[0]));
  const double u2 = 6.0 * ((y[3] - y[2]) / (x[3] - x[2]) - (y[2] - y[1]) / h1);
  const double z1 = (v2 * u1 - h1 * u2) / (v1 * v2 - h1 * h1);
  const double z2 = (v1 * u2 - h1 * u1) / (v1 * v2 - h1 * h1);
  const double pa = t - x[1], pb = x[2] - t;
  return z2 / (6.0 * h1) * (pa * pa * pa) + z1 / (6.0 * h1) * (pb * pb * pb) +
         (y[2] / h1 - z2 / 6.0 * h1) * pa + (y[1] / h1 - z1 / 6.0 * h1) * pb;
}

template <typename T>
__global__ void k_gather(const GatherArgs g, const MarchArgs<T> a) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = blockIdx.y;
  const int nr = g.rx_count[slot];
  if (tid >= nr * g.ncomp_group) return;
  const int r = tid / g.ncomp_group, cg = tid % g.ncomp_group;
  const int comp = g.comps[cg];
  const int ir = g.rx_local[g.rx_first[slot] + r];
  const double x = g.rxx[ir], y = g.rxy[ir], z = g.rxz[ir];
  const int ix = g.rix[ir], iy = g.riy[ir], iz = g.riz[ir];
  const double *X = g.X, *Y = g.Y, *Z = g.Z, *dx = g.dx, *dy = g.dy, *dz = g.dz;
  // centre (0) or edge (1) stencils per axis; edge axes use the half-cell test
  const int xe = (comp == 1 || comp == 2 || comp == 3);
  const int ye = (comp == 0 || comp == 2 || comp == 4);
  const int ze = (comp == 0 || comp == 1 || comp == 5);
  const int s1 = xe ? ((X[ix - 1] + 0.5 * dx[ix - 1]) > x ? ix - 2 : ix - 1) : ix - 2;
  const int s2 = ye ? ((Y[iy - 1] + 0.5 * dy[iy - 1]) > y ? iy - 2 : iy - 1) : iy - 2;
  int s3;
  if (ze) s3 = (iz >= 1 && (Z[iz - 1] + 0.5 * dz[iz - 1]) > z) ? iz - 2 : iz - 1;
  else s3 = iz - 2;
  if (s3 < 0) s3 = 0;
  double tx[4], ty[4], tz[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    tx[q] = xe ? X[s1 + q] - 0.5 * dx[s1 + q] : X[s1 + q];
    ty[q] = ye ? Y[s2 + q] - 0.5 * dy[s2 + q] : Y[s2 + q];
    tz[q] = ze ? Z[s3 + q] - 0.5 * dz[s3 + q] : Z[s3 + q];
  }
  const long long fo = (long long)slot * g.fstride;
  double v;
  double Y1[4], Y2[4], F[4];
  if (comp < 5) {
    const T *f = (const T *)g.f[comp] + fo;
    for (int ik = 0; ik < 4; ++ik) {
      for (int ij = 0; ij < 4; ++ij) {
        const long long o = (long long)(ik + s3) * g.PS + (long long)(ij + s2) * g.PX + s1;
#pragma unroll
        for (int q = 0; q < 4; ++q) F[q] = (double)f[o + q];
        Y1[ij] = spline4(tx, F, x);
      }
      Y2[ik] = spline4(ty, Y1, y);
    }
    v = spline4(tz, Y2, z);
  } else {
    // Bz at levels s3..s3+3 of the 4x4 columns: bottom-up running sums of div_h(Bx, By)
    const T *bxf = (const T *)g.f[3] + fo, *byf = (const T *)g.f[4] + fo;
    double Bzv[4][4][4];  // [ik][ij][q]
    for (int ij = 0; ij < 4; ++ij)
      for (int q = 0; q < 4; ++q) {
        const int ii = s1 + q, jj = s2 + ij;
        const T rdx = a.rdx[ii + 1], rdy = a.rdy[jj + 1];
        Acc<T> acc;
        acc.init();
        for (int ik = 0; ik < 4; ++ik)
          if (s3 + ik >= g.N) Bzv[ik][ij][q] = 0.0;
        for (int k = g.N - 1; k >= s3; --k) {
          const long long o = (long long)k * g.PS + (long long)jj * g.PX + ii;
          acc.add(a.dzt[k] * ((bxf[o + 1] - bxf[o]) * rdx + (byf[o + g.PX] - byf[o]) * rdy));
          if (k <= s3 + 3) Bzv[k - s3][ij][q] = (double)acc.val();
        }
      }
    for (int ik = 0; ik < 4; ++ik) {
      for (int ij = 0; ij < 4; ++ij) Y1[ij] = spline4(tx, Bzv[ik][ij], x);
      Y2[ik] = spline4(ty, Y1, y);
    }
    v = spline4(tz, Y2, z);
  }
  const int c = g.comp_base + cg;
  double *e0 = g.ef0 + (long long)slot * 2 * g.ncomp * g.nrmax;
  e0[(long long)g.slot_store * g.ncomp * g.nrmax + (long long)c * g.nrmax + r] = v;
  double *efs = g.efs + (long long)slot * g.nrmax * g.ncomp * g.nt;
  for (int o = 0; o < g.nops; ++o) {
    const double l1 = g.op_l1[o];
    efs[((long long)r * g.ncomp + c) * g.nt + g.op_it[o]] =
        (1.0 - l1) * e0[(long long)c * g.nrmax + r] + l1 * e0[(long long)g.ncomp * g.nrmax + (long long)c * g.nrmax + r];
  }
}

__global__ void k_fill_zero(uint32_t *p, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = 0u;
}

// ---- launchers -----------------------------------------------------------------------------------
template <typename T>
hipError_t launch_march(const MarchArgs<T> &a, int nslot, hipStream_t s) {
  dim3 blk(kMarchLanes, kMarchRows), grd(8 * a.nper, nslot);
  hipLaunchKernelGGL((k_march<T, kMarchRows>), grd, blk, 0, s, a);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_surface(const MarchArgs<T> &a, int nslot, hipStream_t s) {
  dim3 blk(kMarchLanes, kMarchRows), grd(8 * a.nper, nslot);
  hipLaunchKernelGGL((k_surface<T, kMarchRows>), grd, blk, 0, s, a);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_bz_volume(const MarchArgs<T> &a, T *bz, int slot, hipStream_t s) {
  dim3 blk(64), grd((a.L + 63) / 64, a.M);
  hipLaunchKernelGGL((k_bz_volume<T>), grd, blk, 0, s, a, bz, slot);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_gather(const GatherArgs &g, int is_b, int nslot, const MarchArgs<T> &a, hipStream_t s) {
  (void)is_b;
  const int n = g.nrmax * g.ncomp_group;
  dim3 blk(64), grd((n + 63) / 64, nslot);
  hipLaunchKernelGGL((k_gather<T>), grd, blk, 0, s, g, a);
  return hipGetLastError();
}
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 4;
  if (n == 0) return hipSuccess;
  size_t nb = (n + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)nb), dim3(256), 0, s, (uint32_t *)p, n);
  return hipGetLastError();
}

template hipError_t launch_march<float>(const MarchArgs<float> &, int, hipStream_t);
template hipError_t launch_march<double>(const MarchArgs<double> &, int, hipStream_t);
template hipError_t launch_surface<float>(const MarchArgs<float> &, int, hipStream_t);
template hipError_t launch_surface<double>(const MarchArgs<double> &, int, hipStream_t);
template hipError_t launch_bz_volume<float>(const MarchArgs<float> &, float *, int, hipStream_t);
template hipError_t launch_bz_volume<double>(const MarchArgs<double> &, double *, int, hipStream_t);
template hipError_t launch_gather<float>(const GatherArgs &, int, int, const MarchArgs<float> &, hipStream_t);
template hipError_t launch_gather<double>(const GatherArgs &, int, int, const MarchArgs<double> &, hipStream_t);

}  // namespace tcsem
