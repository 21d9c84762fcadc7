// tcsem_upcont.hip -- gfx950 kernels of the air boundary: upward continuation of Bz(0) to
// BxAir/ByAir at z = -dz0/2 (UpContinuation_gpu.cu:166-274).
//
// The operator is linear in Bz(0).  Two chains compute it; both end in k_up_bilin (FP64
// arithmetic).  The tiled chain computes in the field precision T (its FFT twiddles, Bluestein
// chirps, window weights and gains are FP64 tables, rounded once on use); the line chain in FP64.
//
// Tiled chain (both FFT working lengths <= 2048), four kernels, every global access coalesced:
//   k_up_rows_fwd  one wavefront per pair of irregular rows: the natural-spline IR->R resampling
//                  along x (SplineGridConvIR2R.cu:539-752, window weights tap-major) of both rows
//                  into one complex line, its FFT along x, and the split into the two real-input
//                  half spectra X(k) = (Z(k) + conj Z(-k)) / 2, (Z(k) - conj Z(-k)) / 2i; the
//                  workgroup's rows leave through LDS as contiguous runs of a kx-major plane;
//   k_up_cols      one wavefront per half-spectrum column kx (contiguous in that plane): the
//                  y-padding rows of the regular plane as window sums of the irregular rows, the
//                  FFT along y, and the gains of FBz_to_FBxFBy (UpContinuation_gpu.cu:113-164)
//                  with the bins that are not Hermitian-consistent zeroed (below); the column
//                  spectrum stays in registers while the two inverse FFTs along y run, and
//                  (Gx, Gy) go out as one contiguous column;
//   k_up_rows_inv  one wavefront per regular row: a tile of rows is read kx-major, each row is
//                  completed to the full spectrum Gx + i Gy by Hermitian symmetry (kx and L1 - kx),
//                  and one inverse FFT along x gives Bx = Re, By = Im.
// Line chain (any length), six kernels:
//   k_up_xpass  one workgroup per irregular row: natural-spline IR->R resampling along x
//               (SplineGridConvIR2R.cu:539-752) as precomputed window weights, tap-major so that
//               lanes over columns read them coalesced; the row is staged in LDS;
//   k_up_ypass  one workgroup per y-padding row of the regular plane: the y pass of the same
//               resampling, lanes over columns, one weight per tap for the whole workgroup (the
//               interior rows are the x-pass rows themselves and are not copied);
//   k_up_ycols  one wavefront per regular column: forward FFT along y in LDS;
//   k_up_xrows  one wavefront per wavenumber row: forward FFT along x, the wavenumber filter of
//               FBz_to_FBxFBy (UpContinuation_gpu.cu:113-164) continued to the full spectrum by
//               Hermitian symmetry, and one inverse FFT along x of Z = FBx + i FBy;
//   k_up_ycols2 one wavefront per column: inverse FFT along y; Re Z = Bx, Im Z = By;
//   k_up_bilin  bilinear regular -> irregular with the 1/(L1*M1) normalisation
//               (UpContinuation_gpu.cu:42-110).
// The reference's C2R transform returns the inverse of the Hermitian part of its half spectrum:
// at a self-conjugate bin it keeps the real part only, and i*g*F is purely imaginary there
// because F is the transform of a real plane.  The bins where the half-spectrum gains are not
// already Hermitian-consistent are: the DC bin, the whole x-Nyquist column of the x gain (even in
// ky, so its Hermitian part is zero) and the y-Nyquist bins of the y gain at kx = 0 and
// kx = L1/2.  With those gains set to zero both spectra are exactly Hermitian, both inverses are
// real, and one complex inverse of FBx + i FBy yields Bx and By at once.  (The tiled chain takes
// the real part at kx = 0 and kx = L1/2 of each row, as a real-output inverse does.)
// FFTs: working lengths up to 2048 run Stockham radix-8/4/2 passes between two LDS buffers with
// the twiddle table staged in LDS, one wavefront per line; longer ones run in-place radix-2 with
// 256 threads per line and the table in global memory.  Lengths that are not powers of two use
// Bluestein's chirp-z algorithm on a power-of-two length.  Tables are FP64 from the host.
#include <mutex>
#include <vector>

#include "tcsem_internal.h"

namespace tcsem {

// Complex arithmetic on float2 / double2 (V), scalar S.
template <typename S>
struct CxOf;
template <>
struct CxOf<float> {
  using V = float2;
};
template <>
struct CxOf<double> {
  using V = double2;
};
template <typename V>
__device__ inline V mkc(decltype(V::x) x, decltype(V::x) y) {
  V v;
  v.x = x;
  v.y = y;
  return v;
}
template <typename V>
__device__ inline V cmul(V a, V b) {
  return mkc<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ inline V cadd(V a, V b) {
  return mkc<V>(a.x + b.x, a.y + b.y);
}
template <typename V>
__device__ inline V csub(V a, V b) {
  return mkc<V>(a.x - b.x, a.y - b.y);
}
template <typename V>
__device__ inline V conj2(V a) {
  return mkc<V>(a.x, -a.y);
}
// a * (s i) for real s: (x + i y) s i = -s y + i s x
template <typename V>
__device__ inline V muli(V a, decltype(V::x) s) {
  return mkc<V>(-s * a.y, s * a.x);
}
// entry k of an FP64 complex table, in the precision of V
template <typename V>
__device__ inline V from_table(const double *t, int k) {
  const double2 w = ((const double2 *)t)[k];
  return mkc<V>((decltype(V::x))w.x, (decltype(V::x))w.y);
}

// exp(sign 2 pi i q / m) from the m-entry table of exp(-2 pi i q / m), 0 <= q < m
template <typename V>
__device__ inline V twid(const V *tw, int q, int sign) {
  V w = tw[q];
  if (sign > 0) w.y = -w.y;
  return w;
}

// in-register DFTs of size R with exp(sign 2 pi i / R) as the root
template <int R, typename V>
struct Dft;
template <typename V>
struct Dft<2, V> {
  __device__ static inline void run(V *v, int) {
    const V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <typename V>
struct Dft<4, V> {
  __device__ static inline void run(V *v, int sign) {
    using S = decltype(V::x);
    const S s = (S)sign;
    const V t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const V t2 = cadd(v[1], v[3]), t3 = muli(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};
template <typename V>
struct Dft<8, V> {
  __device__ static inline void run(V *v, int sign) {
    using S = decltype(V::x);
    const S s = (S)sign, h = (S)0.70710678118654752440;  // cos(pi/4)
    V e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, V>::run(e, sign);
    Dft<4, V>::run(o, sign);
    // w^k, w = exp(sign i pi / 4): w = (h, s h), w^2 = s i, w^3 = (-h, s h)
    o[1] = cmul(o[1], mkc<V>(h, s * h));
    o[2] = muli(o[2], s);
    o[3] = cmul(o[3], mkc<V>(-h, s * h));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

// One Stockham autosort pass of radix R over length m, reading a and writing b; Ns is the
// length of the sub-transforms already done.  Butterfly j reads a[j + r m/R], applies the
// twiddles exp(sign 2 pi i r (j mod Ns) / (Ns R)), and writes b[(j - j mod Ns) R + j mod Ns + r Ns].
template <int R, typename V>
__device__ inline void stockham_pass(const V *a, V *b, int m, int Ns, const V *tw, int sign, int tid, int nthr) {
  const int nb = m / R, span = m / (Ns * R);
  for (int j = tid; j < nb; j += nthr) {
    const int k = j & (Ns - 1);
    V v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = a[j + r * nb];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid(tw, r * k * span, sign));
    }
    Dft<R, V>::run(v, sign);
    const int d = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) b[d + r * Ns] = v[r];
  }
}

// Unnormalised DFT of length m = 2^log2m of the line in a; with PP the passes alternate between
// a and b (m entries each), otherwise they run in place in a.  Returns the buffer holding the
// result.  Every thread of the workgroup calls it with the same m (it synchronises the workgroup).
template <bool PP, typename V>
__device__ V *fft_pow2(V *a, V *b, int m, int log2m, const V *tw, int sign, int tid, int nthr) {
  if constexpr (PP) {
    int Ns = 1, left = log2m;
    while (left > 0) {
      if (left >= 3) {
        stockham_pass<8>(a, b, m, Ns, tw, sign, tid, nthr);
        Ns <<= 3;
        left -= 3;
      } else if (left == 2) {
        stockham_pass<4>(a, b, m, Ns, tw, sign, tid, nthr);
        Ns <<= 2;
        left -= 2;
      } else {
        stockham_pass<2>(a, b, m, Ns, tw, sign, tid, nthr);
        Ns <<= 1;
        left -= 1;
      }
      __syncthreads();
      V *t = a;
      a = b;
      b = t;
    }
    return a;
  } else {
    // in-place radix-2 decimation in time
    for (int idx = tid; idx < m; idx += nthr) {
      const int r = (int)(__brev((unsigned)idx) >> (32 - log2m));
      if (idx < r) {
        const V t = a[idx];
        a[idx] = a[r];
        a[r] = t;
      }
    }
    __syncthreads();
    for (int s = 1; s <= log2m; ++s) {
      const int half = 1 << (s - 1), stride = m >> s;
      for (int bf = tid; bf < (m >> 1); bf += nthr) {
        const int grp = bf >> (s - 1), pos = bf & (half - 1);
        const int i0 = (grp << s) + pos, i1 = i0 + half;
        const V u = a[i0], v = cmul(a[i1], twid(tw, pos * stride, sign));
        a[i0] = cadd(u, v);
        a[i1] = csub(u, v);
      }
      __syncthreads();
    }
    return a;
  }
}

// Unnormalised DFT of length P.n of a[0..n) (a and, with PP, b hold P.m entries; tw is the
// m-entry twiddle table in the precision of V, staged in LDS with PP; the Bluestein chirp tables
// are FP64 and rounded on use).  The caller has written a[0..n) and synchronised; returns the
// buffer holding the result in [0, n).
template <bool PP, typename V>
__device__ V *fft_line(V *a, V *b, const FftPlan &P, const V *tw, int sign, int tid, int nthr) {
  if (P.m == P.n) return fft_pow2<PP>(a, b, P.m, P.log2m, tw, sign, tid, nthr);
  for (int k = tid; k < P.m; k += nthr) {
    if (k < P.n) {
      V x = a[k];
      if (sign > 0) x = conj2(x);  // inverse = conj(forward(conj))
      a[k] = cmul(x, from_table<V>(P.chirp, k));
    } else {
      a[k] = mkc<V>(0, 0);
    }
  }
  __syncthreads();
  V *r = fft_pow2<PP>(a, b, P.m, P.log2m, tw, -1, tid, nthr);
  V *o = (r == a) ? b : a;
  for (int k = tid; k < P.m; k += nthr) r[k] = cmul(r[k], from_table<V>(P.bhat, k));
  __syncthreads();
  r = fft_pow2<PP>(r, o, P.m, P.log2m, tw, +1, tid, nthr);
  for (int k = tid; k < P.n; k += nthr) {
    V y = cmul(r[k], from_table<V>(P.chirp, k));
    if (sign > 0) y = conj2(y);
    r[k] = y;
  }
  __syncthreads();
  return r;
}

// LDS (double2 entries) of one line and of the staged twiddle table
__host__ __device__ inline int line_span(int m, bool pp) { return pp ? 2 * m : m; }
__host__ __device__ inline int table_span(int m, bool pp) { return pp ? m : 0; }

// Per-workgroup LDS set-up of a line kernel: the line buffers of this wavefront and the twiddle
// table (copied by the whole workgroup; the caller synchronises before use).
template <bool PP>
__device__ inline const double2 *stage_table(double2 *lds, const FftPlan &P, int lines) {
  if constexpr (PP) {
    double2 *t = lds + (long long)lines * line_span(P.m, true);
    const double2 *g = (const double2 *)P.tw;
    const int flat = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
    for (int q = flat; q < P.m; q += nt) t[q] = g[q];
    return t;
  } else {
    return (const double2 *)P.tw;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_up_xpass(const UpcontArgs<T> u) {
  extern __shared__ double srow[];
  const int j = blockIdx.x, slot = blockIdx.y;
  const T *row = u.bz0 + (long long)slot * u.pstride + (long long)j * u.pz;
  for (int i = threadIdx.x; i < u.L; i += blockDim.x) srow[i] = (double)row[i];
  __syncthreads();
  double *trow = u.tx + (long long)slot * u.wstride + (long long)j * u.L1;
  const int ox = u.Nx + u.AL - kNPad, aL = u.AnLx, aR = u.AnRx;
  for (int c = threadIdx.x; c < u.L1; c += blockDim.x) {
    double v = 0.0;
    if (c >= aL && c < aL + u.Nx) {
      v = srow[u.AL + c - aL];
    } else if (c < aL) {
#pragma unroll 8
      for (int q = 0; q < u.n1; ++q) v += u.wl[(long long)q * aL + c] * srow[q];
    } else {
      const int cr = c - aL - u.Nx;
#pragma unroll 8
      for (int q = 0; q < u.n2; ++q) v += u.wr[(long long)q * aR + cr] * srow[ox + q];
    }
    trow[c] = v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_up_ypass(const UpcontArgs<T> u) {
  const int slot = blockIdx.y;
  int r, src;
  const double *w;
  if ((int)blockIdx.x < u.Any) {
    r = blockIdx.x;
    w = u.wf + (long long)r * u.n3;
    src = 0;
  } else {
    const int rb = blockIdx.x - u.Any;
    r = u.Any + u.Ny + rb;
    w = u.wb + (long long)rb * u.n3;
    src = u.Ny + u.AY - kNPad;
  }
  const double *tx = u.tx + (long long)slot * u.wstride + (long long)src * u.L1;
  double *out = u.ry + (long long)slot * u.wstride + (long long)r * u.L1;
  for (int c = threadIdx.x; c < u.L1; c += blockDim.x) {
    double v = 0.0;
#pragma unroll 8
    for (int q = 0; q < u.n3; ++q) v += w[q] * tx[(long long)q * u.L1 + c];
    out[c] = v;
  }
}

template <typename T, bool PP>
__global__ void __launch_bounds__(256) k_up_ycols(const UpcontArgs<T> u) {
  extern __shared__ double2 lds[];
  const int ln = threadIdx.y, tid = threadIdx.x, nthr = blockDim.x;
  const int c = blockIdx.x * blockDim.y + ln, slot = blockIdx.y;
  const bool live = c < u.L1;
  const int m = u.py.m;
  double2 *a = lds + (long long)ln * line_span(m, PP);
  double2 *b = a + m;
  const double2 *tw = stage_table<PP>(lds, u.py, blockDim.y);
  const double *tx = u.tx + (long long)slot * u.wstride, *ry = u.ry + (long long)slot * u.wstride;
  const int y0 = u.Any, y1 = u.Any + u.Ny, sh = u.AY - u.Any;
  for (int r = tid; r < u.M1; r += nthr) {
    double v = 0.0;
    if (live) v = (r >= y0 && r < y1) ? tx[(long long)(r + sh) * u.L1 + c] : ry[(long long)r * u.L1 + c];
    a[r] = make_double2(v, 0.0);
  }
  __syncthreads();
  double2 *res = fft_line<PP>(a, b, u.py, tw, -1, tid, nthr);
  if (live) {
    double2 *out = u.cplx + (long long)slot * u.wstride;
    for (int r = tid; r < u.M1; r += nthr) out[(long long)r * u.L1 + c] = res[r];
  }
}

template <typename T, bool PP>
__global__ void __launch_bounds__(256) k_up_xrows(const UpcontArgs<T> u) {
  extern __shared__ double2 lds[];
  const int ln = threadIdx.y, tid = threadIdx.x, nthr = blockDim.x;
  const int ky = blockIdx.x * blockDim.y + ln, slot = blockIdx.y;
  const int L1 = u.L1, M1 = u.M1, H = u.H, m = u.px.m;
  const bool live = ky < M1;
  double2 *a = lds + (long long)ln * line_span(m, PP);
  double2 *b = a + m;
  const double2 *tw = stage_table<PP>(lds, u.px, blockDim.y);
  double2 *row = u.cplx + (long long)slot * u.wstride + (long long)(live ? ky : 0) * L1;
  for (int x = tid; x < L1; x += nthr) a[x] = live ? row[x] : make_double2(0.0, 0.0);
  __syncthreads();
  double2 *f = fft_line<PP>(a, b, u.px, tw, -1, tid, nthr);
  double2 *o = (f == a) ? b : a;
  const int kym = live ? (M1 - ky) % M1 : 0;
  const bool evenL = (L1 & 1) == 0, evenM = (M1 & 1) == 0;
  // in place: each element is read and written by the same thread
  for (int kx = tid; kx < L1; kx += nthr) {
    const double2 fv = f[kx];
    double gx = 0.0, gy = 0.0, s = 1.0;
    if (live) {
      if (kx < H) {
        gx = u.gx[(long long)ky * H + kx];
        gy = u.gy[(long long)ky * H + kx];
      } else {  // Hermitian completion: S(k) = conj(S(-k)) = -i g(-k) F(k)
        gx = u.gx[(long long)kym * H + (L1 - kx)];
        gy = u.gy[(long long)kym * H + (L1 - kx)];
        s = -1.0;
      }
      if (ky == 0 && kx == 0) gx = gy = 0.0;
      if (evenL && kx == L1 / 2) gx = 0.0;
      if (evenM && ky == M1 / 2 && (kx == 0 || (evenL && kx == L1 / 2))) gy = 0.0;
    }
    // Z = (i s gx) F + i (i s gy) F
    const double2 sx = muli(fv, s * gx), sy = muli(fv, s * gy);
    f[kx] = make_double2(sx.x - sy.y, sx.y + sy.x);
  }
  __syncthreads();
  double2 *z = fft_line<PP>(f, o, u.px, tw, +1, tid, nthr);
  if (live)
    for (int x = tid; x < L1; x += nthr) row[x] = z[x];
}

template <typename T, bool PP>
__global__ void __launch_bounds__(256) k_up_ycols2(const UpcontArgs<T> u) {
  extern __shared__ double2 lds[];
  const int ln = threadIdx.y, tid = threadIdx.x, nthr = blockDim.x;
  const int c = blockIdx.x * blockDim.y + ln, slot = blockIdx.y;
  const int L1 = u.L1, M1 = u.M1, m = u.py.m;
  const bool live = c < L1;
  double2 *a = lds + (long long)ln * line_span(m, PP);
  double2 *b = a + m;
  const double2 *tw = stage_table<PP>(lds, u.py, blockDim.y);
  const double2 *src = u.cplx + (long long)slot * u.wstride;
  for (int r = tid; r < M1; r += nthr) a[r] = live ? src[(long long)r * L1 + c] : make_double2(0.0, 0.0);
  __syncthreads();
  double2 *z = fft_line<PP>(a, b, u.py, tw, +1, tid, nthr);
  if (live) {
    double *bxr = u.bxr + (long long)slot * u.wstride, *byr = u.byr + (long long)slot * u.wstride;
    for (int r = tid; r < M1; r += nthr) {
      bxr[(long long)r * L1 + c] = z[r].x;
      byr[(long long)r * L1 + c] = z[r].y;
    }
  }
}

// ---- tiled chain (computes in the field precision T) -------------------------------------
// The argument record of a workgroup: the launch's own, or (model batch, GEO) the record of the
// model in grid row blockIdx.y, copied by value so that its fields stay scalar loads.
template <bool GEO, typename U>
__device__ inline U slot_record(const U &u0, const U *geo) {
  if constexpr (GEO) return geo[blockIdx.y];
  else return u0;
}
// LDS of a tiled kernel: `lines` line regions of two buffers of m + 2 complex entries each (the
// two extra entries hold the Nyquist bin of an unpacked half spectrum and shift the lines across
// LDS banks), then the m-entry twiddle table.
__host__ __device__ inline int tile_span(int m) { return m + 2; }
template <typename T>
__host__ __device__ inline size_t tile_lds(int m, int lines) {
  return ((size_t)lines * 2 * tile_span(m) + m) * 2 * sizeof(T);
}

// The twiddle table (FP64 on the device, exp(-2 pi i q / m) from the host) staged in LDS in the
// precision of V: each entry rounded once from its FP64 value.
template <typename V>
__device__ inline const V *stage_twiddles(V *dst, const FftPlan &P) {
  const int flat = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
  for (int q = flat; q < P.m; q += nt) dst[q] = from_table<V>(P.tw, q);
  return dst;
}

// Regular column c of one irregular row staged in LDS (the x pass of the IR->R resampling).
template <typename T>
__device__ inline T resample_x(const UpcontArgs<T> &u, const T *srow, int c) {
  const int aL = u.AnLx;
  if (c >= aL && c < aL + u.Nx) return srow[u.AL + c - aL];
  T v = 0;
  if (c < aL) {
#pragma unroll 8
    for (int q = 0; q < u.n1; ++q) v += (T)u.wl[(long long)q * aL + c] * srow[q];
  } else {
    const int cr = c - aL - u.Nx, ox = u.Nx + u.AL - kNPad;
#pragma unroll 8
    for (int q = 0; q < u.n2; ++q) v += (T)u.wr[(long long)q * u.AnRx + cr] * srow[ox + q];
  }
  return v;
}

// blockDim (W, P), W = 64, 128 or 256 threads per line: line p transforms irregular rows
// j0 = 2 (blockIdx.x P + p) and j0 + 1.
template <typename T, bool GEO>
__global__ void __launch_bounds__(256) k_up_rows_fwd(const UpcontArgs<T> u0, const UpcontArgs<T> *geo) {
  using V = typename CxOf<T>::V;
  extern __shared__ char lds_raw[];
  V *lds = (V *)lds_raw;
  const UpcontArgs<T> u = slot_record<GEO>(u0, geo);
  const int lane = threadIdx.x, W = blockDim.x, p = threadIdx.y, np = blockDim.y, slot = GEO ? 0 : blockIdx.y;
  const int m = u.px.m, span = tile_span(m), L1 = u.L1, L = u.L, M = u.M;
  if (GEO && (int)blockIdx.x * 2 * np >= M) return;  // the slot's rows end before this workgroup
  V *a = lds + (long long)p * 2 * span, *b = a + span;
  const V *tw = stage_twiddles(lds + (long long)np * 2 * span, u.px);
  const int jb = blockIdx.x * 2 * np, j0 = jb + 2 * p;
  // the two rows are staged in the second buffer (L <= m + 2 entries each, checked on the host)
  T *s0 = (T *)b, *s1 = s0 + L;
  const T *plane = u.bz0 + (long long)slot * u.pstride;
  for (int i = lane; i < L; i += W) {
    s0[i] = j0 < M ? plane[(long long)j0 * u.pz + i] : (T)0;
    s1[i] = j0 + 1 < M ? plane[(long long)(j0 + 1) * u.pz + i] : (T)0;
  }
  __syncthreads();
  for (int c = lane; c < L1; c += W) a[c] = mkc<V>(resample_x(u, s0, c), resample_x(u, s1, c));
  __syncthreads();
  V *f = fft_line<true>(a, b, u.px, tw, -1, lane, W);
  V *o = (f == a) ? b : a;
  const T h = (T)0.5;
  for (int k = lane; k < u.H; k += W) {
    const V z = f[k], w = f[k == 0 ? 0 : L1 - k];
    o[2 * k] = mkc<V>(h * (z.x + w.x), h * (z.y - w.y));       // (Z(k) + conj Z(-k)) / 2
    o[2 * k + 1] = mkc<V>(h * (z.y + w.y), -h * (z.x - w.x));  // (Z(k) - conj Z(-k)) / 2i
  }
  __syncthreads();
  // out through LDS: for each kx the workgroup's 2 np rows are one contiguous run
  const int R = 2 * np, ooff = (f == a) ? span : 0;
  V *A = (V *)u.A + (long long)slot * u.sA;
  for (int idx = p * W + lane; idx < u.H * R; idx += np * W) {
    const int kx = idx / R, r = idx - kx * R, j = jb + r;
    if (j < M) A[(long long)kx * u.MA + j] = lds[(long long)(r >> 1) * 2 * span + ooff + 2 * kx + (r & 1)];
  }
}

// blockDim (W, Q), W = 64, 128 or 256 threads per line: line q owns half-spectrum column
// kx = blockIdx.x Q + q.  E >= M1 / W: column entries kept in registers per thread.
template <typename T, int E, bool GEO>
__global__ void __launch_bounds__(256) k_up_cols(const UpcontArgs<T> u0, const UpcontArgs<T> *geo) {
  using V = typename CxOf<T>::V;
  extern __shared__ char lds_raw[];
  V *lds = (V *)lds_raw;
  const UpcontArgs<T> u = slot_record<GEO>(u0, geo);
  const int lane = threadIdx.x, W = blockDim.x, q = threadIdx.y, nq = blockDim.y, slot = GEO ? 0 : blockIdx.y;
  const int m = u.py.m, span = tile_span(m), M1 = u.M1, M = u.M;
  if (GEO && (int)blockIdx.x * nq >= u.H) return;
  V *a = lds + (long long)q * 2 * span, *b = a + span;
  const V *tw = stage_twiddles(lds + (long long)nq * 2 * span, u.py);
  const int kx = blockIdx.x * nq + q;
  const bool live = kx < u.H;
  const int kc = live ? kx : 0;
  const V *Acol = (const V *)u.A + (long long)slot * u.sA + (long long)kc * u.MA;
  for (int j = lane; j < M; j += W) b[j] = live ? Acol[j] : mkc<V>(0, 0);
  __syncthreads();
  // the regular column: core rows are irregular rows, padding rows the y-window sums
  // (the y pass of SplineGridConvIR2R.cu:539-752)
  const int y0 = u.Any, y1 = u.Any + u.Ny, sh = u.AY - u.Any, ob = u.Ny + u.AY - kNPad;
  for (int r = lane; r < M1; r += W) {
    V v;
    if (r >= y0 && r < y1) {
      v = b[r + sh];
    } else {
      const double *w = r < y0 ? u.wf + (long long)r * u.n3 : u.wb + (long long)(r - y1) * u.n3;
      const V *s = r < y0 ? b : b + ob;
      T vx = 0, vy = 0;
#pragma unroll 8
      for (int t = 0; t < u.n3; ++t) {
        const T wt = (T)w[t];
        vx += wt * s[t].x;
        vy += wt * s[t].y;
      }
      v = mkc<V>(vx, vy);
    }
    a[r] = v;
  }
  __syncthreads();
  V *f = fft_line<true>(a, b, u.py, tw, -1, lane, W);
  V *o = (f == a) ? b : a;
  const double *gx = u.gxT + (long long)kc * M1, *gy = u.gyT + (long long)kc * M1;
  V F[E], Gx[E];
  // F stays in registers; each thread rewrites only its own entries
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int ky = lane + W * e;
    F[e] = ky < M1 ? f[ky] : mkc<V>(0, 0);
    if (ky < M1) f[ky] = muli(F[e], (T)gx[ky]);  // FBx = i gx F
  }
  __syncthreads();
  V *rx = fft_line<true>(f, o, u.py, tw, +1, lane, W);
  V *o2 = (rx == f) ? o : f;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int ky = lane + W * e;
    Gx[e] = ky < M1 ? rx[ky] : mkc<V>(0, 0);
    if (ky < M1) rx[ky] = muli(F[e], (T)gy[ky]);  // FBy = i gy F
  }
  __syncthreads();
  V *ry = fft_line<true>(rx, o2, u.py, tw, +1, lane, W);
  if (live) {
    V *Gc = (V *)u.G + (long long)slot * u.sG + (long long)kx * u.MG * 2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int y = lane + W * e;
      if (y < M1) {
        Gc[2 * y] = Gx[e];
        Gc[2 * y + 1] = ry[y];
      }
    }
  }
}

// blockDim (W, Q), W = 64, 128 or 256 threads per line: line q inverts regular row
// y = yrows[blockIdx.x Q + q] along x (only the rows the bilinear step reads).
template <typename T, bool GEO>
__global__ void __launch_bounds__(256) k_up_rows_inv(const UpcontArgs<T> u0, const UpcontArgs<T> *geo) {
  using V = typename CxOf<T>::V;
  extern __shared__ char lds_raw[];
  V *lds = (V *)lds_raw;
  const UpcontArgs<T> u = slot_record<GEO>(u0, geo);
  const int lane = threadIdx.x, W = blockDim.x, q = threadIdx.y, nq = blockDim.y, slot = GEO ? 0 : blockIdx.y;
  const int m = u.px.m, span = tile_span(m), L1 = u.L1, H = u.H, nyr = u.nyr;
  if (GEO && (int)blockIdx.x * nq >= nyr) return;
  const V *tw = stage_twiddles(lds + (long long)nq * 2 * span, u.px);
  const int yb = blockIdx.x * nq;  // first entry of the workgroup in the row list
  const V *G = (const V *)u.G + (long long)slot * u.sG;
  const bool evenL = (L1 & 1) == 0;
  // tile in: for each kx the nq rows are one contiguous run of (Gx, Gy) pairs
  for (int idx = q * W + lane; idx < H * nq; idx += nq * W) {
    const int kx = idx / nq, r = idx - kx * nq, y = yb + r < nyr ? u.yrows[yb + r] : -1;
    V gx = mkc<V>(0, 0), gy = gx;
    if (y >= 0) {
      gx = G[((long long)kx * u.MG + y) * 2];
      gy = G[((long long)kx * u.MG + y) * 2 + 1];
    }
    V *row = lds + (long long)r * 2 * span;
    if (kx == 0 || (evenL && 2 * kx == L1)) {
      row[kx] = mkc<V>(gx.x, gy.x);  // self-conjugate bins: real parts
    } else {
      row[kx] = mkc<V>(gx.x - gy.y, gx.y + gy.x);       // Gx + i Gy
      row[L1 - kx] = mkc<V>(gx.x + gy.y, gy.x - gx.y);  // conj Gx + i conj Gy
    }
  }
  __syncthreads();
  V *a = lds + (long long)q * 2 * span, *b = a + span;
  V *z = fft_line<true>(a, b, u.px, tw, +1, lane, W);
  if (yb + q < nyr) {
    const int y = u.yrows[yb + q];
    T *bx = u.tbx + (long long)slot * u.sX + (long long)y * u.XR;
    T *by = u.tby + (long long)slot * u.sX + (long long)y * u.XR;
    for (int x = lane; x < L1; x += W) {
      const V v = z[x];
      bx[x] = v.x;
      by[x] = v.y;
    }
  }
}

// Bilinear R -> IR (UpContinuation_gpu.cu:42-110) from the regular results fx, fy (rows of
// pitch XR, slot stride sX; FP64 from the line chain, the field type from the tiled chain),
// arithmetic in FP64.
template <typename T, typename B, bool GEO>
__global__ void k_up_bilin(const UpcontArgs<T> u0, const B *fx0, const B *fy0, const UpcontArgs<T> *geo) {
  const UpcontArgs<T> u = slot_record<GEO>(u0, geo);
  const B *fx = GEO ? (const B *)u.tbx : fx0, *fy = GEO ? (const B *)u.tby : fy0;
  const int t = blockIdx.x * blockDim.x + threadIdx.x, slot = GEO ? 0 : blockIdx.y;
  const int L0 = u.L, M0 = u.M, L1 = u.L1, M1 = u.M1, XR = u.XR;
  const double norm = (double)u.L1 * (double)M1;
  const long long wo = (long long)slot * u.sX, po = (long long)slot * u.pstride;
  const int nbx = M0 * (L0 + 1);
  if (t < nbx) {
    const int iy = t / (L0 + 1), ix = t % (L0 + 1);
    const int II = min(max(u.xBx2N[ix], 1), L1 - 1), JJ = min(max(u.yo2N[iy], 1), M1 - 1);
    const double Xp = (ix == L0) ? u.X[L0 - 1] + 0.5 * u.dx[L0 - 1] : u.X[ix] - 0.5 * u.dx[ix];
    const B *F = fx + wo;
    double l1 = (Xp - u.Xn[II - 1]) / (u.Xn[II] - u.Xn[II - 1]);
    const double a0 = (1.0 - l1) * (double)F[(long long)(JJ - 1) * XR + II - 1] + l1 * (double)F[(long long)(JJ - 1) * XR + II];
    const double a1 = (1.0 - l1) * (double)F[(long long)JJ * XR + II - 1] + l1 * (double)F[(long long)JJ * XR + II];
    l1 = (u.Y[iy] - u.Yn[JJ - 1]) / (u.Yn[JJ] - u.Yn[JJ - 1]);
    u.bxair[po + (long long)iy * u.pxa + ix] = (T)(((1.0 - l1) * a0 + l1 * a1) / norm);
  } else if (t < nbx + (M0 + 1) * L0) {
    const int tt = t - nbx;
    const int iy = tt / L0, ix = tt % L0;
    const int II = min(max(u.xo2N[ix], 1), L1 - 1), JJ = min(max(u.yBy2N[iy], 1), M1 - 1);
    const double Yp = (iy == M0) ? u.Y[M0 - 1] + 0.5 * u.dy[M0 - 1] : u.Y[iy] - 0.5 * u.dy[iy];
    const B *F = fy + wo;
    double l1 = (u.X[ix] - u.Xn[II - 1]) / (u.Xn[II] - u.Xn[II - 1]);
    const double a0 = (1.0 - l1) * (double)F[(long long)(JJ - 1) * XR + II - 1] + l1 * (double)F[(long long)(JJ - 1) * XR + II];
    const double a1 = (1.0 - l1) * (double)F[(long long)JJ * XR + II - 1] + l1 * (double)F[(long long)JJ * XR + II];
    l1 = (Yp - u.Yn[JJ - 1]) / (u.Yn[JJ] - u.Yn[JJ - 1]);
    u.byair[po + (long long)iy * u.pya + ix] = (T)(((1.0 - l1) * a0 + l1 * a1) / norm);
  }
}

namespace {
// Launch shape of a line kernel: one wavefront per line (4 lines per workgroup when there are
// enough lines to fill the chip) for Stockham lengths; one 256-thread workgroup per line for the
// in-place lengths.
struct LineShape {
  int lines, threads;
  size_t lds;
};
LineShape line_shape(const FftPlan &P, bool pp, long long nlines) {
  const size_t per = (size_t)line_span(P.m, pp) * sizeof(double2), tab = (size_t)table_span(P.m, pp) * sizeof(double2);
  if (!pp) return {1, 256, per};
  const int lines = (nlines >= 2048 && per * 4 + tab <= 96 * 1024) ? 4 : 1;
  return {lines, 64, per * lines + tab};
}
template <typename K>
hipError_t allow_lds(K kern, size_t bytes) {
  constexpr size_t kDefaultLds = 64 * 1024;
  if (bytes <= kDefaultLds) return hipSuccess;
  return hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
template <typename K, typename A>
hipError_t launch_lines(K kern, const LineShape &sh, int nlines, int nslot, hipStream_t s, const A &u) {
  hipError_t e = allow_lds(kern, sh.lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((nlines + sh.lines - 1) / sh.lines, nslot), dim3(sh.threads, sh.lines), sh.lds, s, u);
  return hipGetLastError();
}

// Launch shape of a tiled kernel: with enough lines for several waves per SIMD, one wavefront per
// line and 4 (then 2) lines per workgroup while two such workgroups fit a CU's LDS; with few lines
// (a single slot), one line per workgroup of 256 threads, so that the resampling sums, the
// butterflies and the transfers of a line spread over four wavefronts.
struct Tile {
  int w, q;
  size_t lds;
};
// Threads per line and lines per workgroup.  A chain that shares the CUs with a march takes one
// wavefront per line once there are many lines, so that it holds few CU slots; a chain with
// nothing beside it (step A's chain in a pair of steps) takes 256 threads per line, whose
// shorter per-line latency wins when it has the GPU to itself.  Each line's arithmetic does not
// depend on the shape.
template <typename T>
Tile tile_shape(int m, long long nlines, bool alone) {
  constexpr size_t kHalfCu = 80 * 1024;
  if (alone) return {256, 1, tile_lds<T>(m, 1)};
  if (nlines >= 1024) {
    if (tile_lds<T>(m, 4) <= kHalfCu) return {64, 4, tile_lds<T>(m, 4)};
    if (tile_lds<T>(m, 2) <= kHalfCu) return {64, 2, tile_lds<T>(m, 2)};
    return {64, 1, tile_lds<T>(m, 1)};
  }
  return {256, 1, tile_lds<T>(m, 1)};
}

// geo (model batch): u carries the largest FFT lengths, M, M1 and H over the slots' records, so
// the grid, the tile shape and the LDS cover every slot; workgroups beyond a slot's own lines
// return at once.
template <typename T, int E, bool GEO>
hipError_t launch_cols_e(const UpcontArgs<T> &u, const Tile &t, int nslot, hipStream_t s, const UpcontArgs<T> *geo) {
  hipError_t e = allow_lds(k_up_cols<T, E, GEO>, t.lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_up_cols<T, E, GEO>), dim3((u.H + t.q - 1) / t.q, nslot), dim3(t.w, t.q), t.lds, s, u, geo);
  return hipGetLastError();
}

template <typename T, bool GEO>
hipError_t launch_cols(const UpcontArgs<T> &u, int nslot, hipStream_t s, const UpcontArgs<T> *geo, bool alone) {
  const Tile t = tile_shape<T>(u.py.m, (long long)u.H * nslot, alone);
  const int need = (u.M1 + t.w - 1) / t.w;  // column entries per thread
  if (need <= 1) return launch_cols_e<T, 1, GEO>(u, t, nslot, s, geo);
  if (need <= 2) return launch_cols_e<T, 2, GEO>(u, t, nslot, s, geo);
  if (need <= 4) return launch_cols_e<T, 4, GEO>(u, t, nslot, s, geo);
  if (need <= 8) return launch_cols_e<T, 8, GEO>(u, t, nslot, s, geo);
  if (need <= 16) return launch_cols_e<T, 16, GEO>(u, t, nslot, s, geo);
  return launch_cols_e<T, 32, GEO>(u, t, nslot, s, geo);
}

template <typename T, bool GEO>
hipError_t launch_tiled(const UpcontArgs<T> &u, int nslot, hipStream_t s, hipEvent_t e0, const UpcontArgs<T> *geo,
                        bool alone) {
  const int npair = (u.M + 1) / 2;
  Tile t = tile_shape<T>(u.px.m, (long long)npair * nslot, alone);
  hipError_t e = allow_lds(k_up_rows_fwd<T, GEO>, t.lds);
  if (e != hipSuccess) return e;
  launch_k(k_up_rows_fwd<T, GEO>, dim3((npair + t.q - 1) / t.q, nslot), dim3(t.w, t.q), (std::uint32_t)t.lds, s, e0,
           nullptr, u, geo);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_cols<T, GEO>(u, nslot, s, geo, alone)) != hipSuccess) return e;
  t = tile_shape<T>(u.px.m, (long long)u.nyr * nslot, alone);
  if ((e = allow_lds(k_up_rows_inv<T, GEO>, t.lds)) != hipSuccess) return e;
  hipLaunchKernelGGL((k_up_rows_inv<T, GEO>), dim3((u.nyr + t.q - 1) / t.q, nslot), dim3(t.w, t.q), t.lds, s, u, geo);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lined(const UpcontArgs<T> &u, int nslot, hipStream_t s, hipEvent_t e0) {
  hipError_t e;
  launch_k(k_up_xpass<T>, dim3(u.M, nslot), dim3(256), (std::uint32_t)(u.L * sizeof(double)), s, e0, nullptr, u);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int npad = u.M1 - u.Ny;
  if (npad > 0) {
    hipLaunchKernelGGL((k_up_ypass<T>), dim3(npad, nslot), dim3(256), 0, s, u);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const bool ppx = u.px.m <= 2048, ppy = u.py.m <= 2048;
  const LineShape sc = line_shape(u.py, ppy, (long long)u.L1 * nslot);
  const LineShape sr = line_shape(u.px, ppx, (long long)u.M1 * nslot);
  e = ppy ? launch_lines(k_up_ycols<T, true>, sc, u.L1, nslot, s, u)
          : launch_lines(k_up_ycols<T, false>, sc, u.L1, nslot, s, u);
  if (e != hipSuccess) return e;
  e = ppx ? launch_lines(k_up_xrows<T, true>, sr, u.M1, nslot, s, u)
          : launch_lines(k_up_xrows<T, false>, sr, u.M1, nslot, s, u);
  if (e != hipSuccess) return e;
  e = ppy ? launch_lines(k_up_ycols2<T, true>, sc, u.L1, nslot, s, u)
          : launch_lines(k_up_ycols2<T, false>, sc, u.L1, nslot, s, u);
  return e;
}
}  // namespace

template <typename T>
hipError_t launch_upcont(const UpcontArgs<T> &u, int nslot, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                         const UpcontArgs<T> *geo, bool alone) {
  if (geo && !u.tiled) return hipErrorInvalidValue;  // model batches run the tiled chain only
  hipError_t e = geo ? launch_tiled<T, true>(u, nslot, s, e0, geo, alone)
                     : (u.tiled ? launch_tiled<T, false>(u, nslot, s, e0, nullptr, alone)
                                : launch_lined<T>(u, nslot, s, e0));
  if (e != hipSuccess) return e;
  const int n = u.M * (u.L + 1) + (u.M + 1) * u.L;
  const dim3 grid((n + 255) / 256, nslot), block(256);
  if (geo)
    launch_k(k_up_bilin<T, T, true>, grid, block, 0u, s, nullptr, e1, u, (const T *)nullptr, (const T *)nullptr, geo);
  else if (u.tiled)
    launch_k(k_up_bilin<T, T, false>, grid, block, 0u, s, nullptr, e1, u, (const T *)u.tbx, (const T *)u.tby,
             (const UpcontArgs<T> *)nullptr);
  else
    launch_k(k_up_bilin<T, double, false>, grid, block, 0u, s, nullptr, e1, u, (const double *)u.bxr,
             (const double *)u.byr, (const UpcontArgs<T> *)nullptr);
  return hipGetLastError();
}

template hipError_t launch_upcont<float>(const UpcontArgs<float> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                         const UpcontArgs<float> *, bool);
template hipError_t launch_upcont<double>(const UpcontArgs<double> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                          const UpcontArgs<double> *, bool);

// ---- work planes of the allocation-free upward continuation ------------------------------------
// The tiled chain's work planes (A, G and the regular-plane Bx, By of one slot) for callers that
// cannot allocate live in this file-scope device array; every device that runs the library holds
// its own copy until the process ends.  In host code the array's name is not a device address:
// its per-device address comes from hipGetSymbolAddress, looked up once per device by
// upcont_scratch_init (tcsem_create calls it with the context's device current), and only that
// address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kUpcontScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<void *> g_scratch;  // device address of `scratch`, per device id
}  // namespace

hipError_t upcont_scratch_init(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0) return hipErrorInvalidDevice;
  if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1, nullptr);
  if (g_scratch[dev]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[dev] = p;
  return hipSuccess;
}

void *upcont_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return dev >= 0 && dev < (int)g_scratch.size() ? g_scratch[dev] : nullptr;
}

}  // namespace tcsem
