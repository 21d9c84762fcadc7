// tcsem_upcont.hip -- gfx950 kernels of the air boundary: upward continuation of Bz(0) to
// BxAir/ByAir at z = -dz0/2 (UpContinuation_gpu.cu:166-274).
//
// The chain is linear in Bz(0) and runs in FP64 whatever the field precision:
//   k_up_cols   one workgroup per regular-grid column: natural-spline IR->R resampling
//               (SplineGridConvIR2R.cu:539-752) as precomputed window weights -- x pass for the
//               column, then y pass -- and the forward FFT along y in LDS;
//   k_up_rows   one workgroup per wavenumber row: forward FFT along x, the wavenumber filter of
//               FBz_to_FBxFBy (UpContinuation_gpu.cu:113-164) continued to the full spectrum by
//               Hermitian symmetry (so the C2R of the reference becomes Re of a complex
//               inverse), and the two inverse FFTs along x;
//   k_up_cols2  one workgroup per column: inverse FFTs along y, real parts;
//   k_up_bilin  bilinear regular -> irregular with the 1/(L1*M1) normalisation
//               (UpContinuation_gpu.cu:42-110).
// FFTs are radix-2 in LDS; other lengths use Bluestein's chirp-z algorithm on a power-of-two
// length, with FP64 twiddle and chirp tables built on the host.
#include "tcsem_internal.h"

namespace tcsem {

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// In-place radix-2 decimation-in-time FFT of buf[0..m), m = 2^log2m, sign -1 forward.
__device__ void radix2(double2 *buf, int m, int log2m, const double2 *tw, int sign) {
  const int nt = blockDim.x, tid = threadIdx.x;
  for (int idx = tid; idx < m; idx += nt) {
    const int r = (int)(__brev((unsigned)idx) >> (32 - log2m));
    if (idx < r) {
      const double2 t = buf[idx];
      buf[idx] = buf[r];
      buf[r] = t;
    }
  }
  __syncthreads();
  for (int s = 1; s <= log2m; ++s) {
    const int half = 1 << (s - 1);
    const int stride = m >> s;  // twiddle index step
    for (int b = tid; b < (m >> 1); b += nt) {
      const int grp = b >> (s - 1), pos = b & (half - 1);
      const int i0 = (grp << s) + pos, i1 = i0 + half;
      double2 w = tw[pos * stride];
      if (sign > 0) w.y = -w.y;
      const double2 u = buf[i0], v = cmul(buf[i1], w);
      buf[i0] = make_double2(u.x + v.x, u.y + v.y);
      buf[i1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
  }
}

// Unnormalised DFT of buf[0..n) in place (buf has room for P.m entries).  The caller has
// written buf[0..n) and synchronised.
__device__ void fft_line(double2 *buf, const FftPlan &P, int sign) {
  const double2 *tw = (const double2 *)P.tw;
  if (P.m == P.n) {
    radix2(buf, P.m, P.log2m, tw, sign);
    return;
  }
  const double2 *w = (const double2 *)P.chirp, *bh = (const double2 *)P.bhat;
  const int nt = blockDim.x, tid = threadIdx.x;
  for (int k = tid; k < P.m; k += nt) {
    if (k < P.n) {
      double2 x = buf[k];
      if (sign > 0) x = conj2(x);  // inverse = conj(forward(conj))
      buf[k] = cmul(x, w[k]);
    } else {
      buf[k] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  radix2(buf, P.m, P.log2m, tw, -1);
  for (int k = tid; k < P.m; k += nt) buf[k] = cmul(buf[k], bh[k]);
  __syncthreads();
  radix2(buf, P.m, P.log2m, tw, +1);
  for (int k = tid; k < P.n; k += nt) {
    double2 y = cmul(buf[k], w[k]);
    if (sign > 0) y = conj2(y);
    buf[k] = y;
  }
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(256) k_up_cols(const UpcontArgs<T> u) {
  extern __shared__ double2 lds[];
  const int c = blockIdx.x, slot = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  double2 *buf = lds;                        // FFT buffer, u.py.m entries
  double *tcol = (double *)(lds + u.py.m);   // x-pass column, M entries
  const T *bz0 = u.bz0 + (long long)slot * u.pstride;
  const int M = u.M, PX = u.PX;
  // x pass for this regular column: T(j, c) for every old row j
  for (int j = tid; j < M; j += nt) {
    const T *row = bz0 + (long long)j * PX;
    double v;
    if (c >= u.AnLx && c < u.AnLx + u.Nx) {
      v = (double)row[u.AL + c - u.AnLx];
    } else if (c < u.AnLx) {
      const double *w = u.wl + (long long)c * u.n1;
      v = 0.0;
      for (int q = 0; q < u.n1; ++q) v += w[q] * (double)row[q];
    } else {
      const int cr = c - u.AnLx - u.Nx, ox = u.Nx + u.AL - kNPad;
      const double *w = u.wr + (long long)cr * u.n2;
      v = 0.0;
      for (int q = 0; q < u.n2; ++q) v += w[q] * (double)row[ox + q];
    }
    tcol[j] = v;
  }
  __syncthreads();
  // y pass: BzAir(r, c) for every regular row r
  const int oy = u.Ny + u.AY - kNPad;
  for (int r = tid; r < u.M1; r += nt) {
    double v;
    if (r >= u.Any && r < u.Any + u.Ny) {
      v = tcol[u.AY + r - u.Any];
    } else if (r < u.Any) {
      const double *w = u.wf + (long long)r * u.n3;
      v = 0.0;
      for (int q = 0; q < u.n3; ++q) v += w[q] * tcol[q];
    } else {
      const double *w = u.wb + (long long)(r - u.Any - u.Ny) * u.n3;
      v = 0.0;
      for (int q = 0; q < u.n3; ++q) v += w[q] * tcol[oy + q];
    }
    buf[r] = make_double2(v, 0.0);
  }
  __syncthreads();
  fft_line(buf, u.py, -1);
  double2 *out = u.cplx + (long long)slot * u.wstride;
  for (int r = tid; r < u.M1; r += nt) out[(long long)r * u.L1 + c] = buf[r];
}

template <typename T>
__global__ void __launch_bounds__(256) k_up_rows(const UpcontArgs<T> u) {
  extern __shared__ double2 lds[];
  const int ky = blockIdx.x, slot = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  const int L1 = u.L1, M1 = u.M1, H = u.H;
  double2 *row = u.cplx + (long long)slot * u.wstride + (long long)ky * L1;
  double2 *row2 = u.cplx2 + (long long)slot * u.wstride + (long long)ky * L1;
  double2 *buf = lds;
  for (int x = tid; x < L1; x += nt) buf[x] = row[x];
  __syncthreads();
  fft_line(buf, u.px, -1);
  // filter: FBx = i g_x FBz on the half spectrum; the Hermitian completion of the half spectrum
  // gives -i g(L1-kx, -ky) FBz on the other half.
  const int kym = (M1 - ky) % M1;
  for (int kx = tid; kx < L1; kx += nt) {
    const double2 f = buf[kx];
    double gx, gy, sgn;
    if (kx < H) {
      gx = u.gx[(long long)ky * H + kx];
      gy = u.gy[(long long)ky * H + kx];
      sgn = 1.0;
    } else {
      gx = u.gx[(long long)kym * H + (L1 - kx)];
      gy = u.gy[(long long)kym * H + (L1 - kx)];
      sgn = -1.0;
    }
    // (i sgn g) * f
    row2[kx] = make_double2(-sgn * gy * f.y, sgn * gy * f.x);
    buf[kx] = make_double2(-sgn * gx * f.y, sgn * gx * f.x);
  }
  __syncthreads();
  fft_line(buf, u.px, +1);
  for (int x = tid; x < L1; x += nt) row[x] = buf[x];
  __syncthreads();
  for (int x = tid; x < L1; x += nt) buf[x] = row2[x];
  __syncthreads();
  fft_line(buf, u.px, +1);
  for (int x = tid; x < L1; x += nt) row2[x] = buf[x];
}

template <typename T>
__global__ void __launch_bounds__(256) k_up_cols2(const UpcontArgs<T> u) {
  extern __shared__ double2 lds[];
  const int c = blockIdx.x, slot = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  const int L1 = u.L1, M1 = u.M1;
  double2 *buf = lds;
  for (int pass = 0; pass < 2; ++pass) {
    const double2 *src = (pass == 0 ? u.cplx : u.cplx2) + (long long)slot * u.wstride;
    double *dst = (pass == 0 ? u.bxr : u.byr) + (long long)slot * u.wstride;
    for (int r = tid; r < M1; r += nt) buf[r] = src[(long long)r * L1 + c];
    __syncthreads();
    fft_line(buf, u.py, +1);
    for (int r = tid; r < M1; r += nt) dst[(long long)r * L1 + c] = buf[r].x;
    __syncthreads();
  }
}

template <typename T>
__global__ void k_up_bilin(const UpcontArgs<T> u) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, slot = blockIdx.y;
  const int L0 = u.L, M0 = u.M, L1 = u.L1, M1 = u.M1;
  const double norm = (double)(L1 * M1);
  const long long wo = (long long)slot * u.wstride, po = (long long)slot * u.pstride;
  const int nbx = M0 * (L0 + 1);
  if (t < nbx) {
    const int iy = t / (L0 + 1), ix = t % (L0 + 1);
    const int II = min(max(u.xBx2N[ix], 1), L1 - 1), JJ = min(max(u.yo2N[iy], 1), M1 - 1);
    const double Xp = (ix == L0) ? u.X[L0 - 1] + 0.5 * u.dx[L0 - 1] : u.X[ix] - 0.5 * u.dx[ix];
    const double *F = u.bxr + wo;
    double l1 = (Xp - u.Xn[II - 1]) / (u.Xn[II] - u.Xn[II - 1]);
    const double a0 = (1.0 - l1) * F[(long long)(JJ - 1) * L1 + II - 1] + l1 * F[(long long)(JJ - 1) * L1 + II];
    const double a1 = (1.0 - l1) * F[(long long)JJ * L1 + II - 1] + l1 * F[(long long)JJ * L1 + II];
    l1 = (u.Y[iy] - u.Yn[JJ - 1]) / (u.Yn[JJ] - u.Yn[JJ - 1]);
    u.bxair[po + (long long)iy * u.PX + ix] = (T)(((1.0 - l1) * a0 + l1 * a1) / norm);
  } else if (t < nbx + (M0 + 1) * L0) {
    const int tt = t - nbx;
    const int iy = tt / L0, ix = tt % L0;
    const int II = min(max(u.xo2N[ix], 1), L1 - 1), JJ = min(max(u.yBy2N[iy], 1), M1 - 1);
    const double Yp = (iy == M0) ? u.Y[M0 - 1] + 0.5 * u.dy[M0 - 1] : u.Y[iy] - 0.5 * u.dy[iy];
    const double *F = u.byr + wo;
    double l1 = (u.X[ix] - u.Xn[II - 1]) / (u.Xn[II] - u.Xn[II - 1]);
    const double a0 = (1.0 - l1) * F[(long long)(JJ - 1) * L1 + II - 1] + l1 * F[(long long)(JJ - 1) * L1 + II];
    const double a1 = (1.0 - l1) * F[(long long)JJ * L1 + II - 1] + l1 * F[(long long)JJ * L1 + II];
    l1 = (Yp - u.Yn[JJ - 1]) / (u.Yn[JJ] - u.Yn[JJ - 1]);
    u.byair[po + (long long)iy * u.PX + ix] = (T)(((1.0 - l1) * a0 + l1 * a1) / norm);
  }
}

template <typename T>
hipError_t launch_upcont(const UpcontArgs<T> &u, int nslot, hipStream_t s) {
  const size_t lds1 = sizeof(double2) * u.py.m + sizeof(double) * u.M;
  constexpr size_t kDefaultLds = 64 * 1024;
  if (lds1 > kDefaultLds || sizeof(double2) * u.px.m > kDefaultLds || sizeof(double2) * u.py.m > kDefaultLds) {
    // planes past 4096 points per axis use up to 160 KiB of LDS per workgroup on gfx950
    (void)hipFuncSetAttribute((const void *)k_up_cols<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
    (void)hipFuncSetAttribute((const void *)k_up_rows<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)(sizeof(double2) * u.px.m));
    (void)hipFuncSetAttribute((const void *)k_up_cols2<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)(sizeof(double2) * u.py.m));
  }
  hipLaunchKernelGGL((k_up_cols<T>), dim3(u.L1, nslot), dim3(256), lds1, s, u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_up_rows<T>), dim3(u.M1, nslot), dim3(256), sizeof(double2) * u.px.m, s, u);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((k_up_cols2<T>), dim3(u.L1, nslot), dim3(256), sizeof(double2) * u.py.m, s, u);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int n = u.M * (u.L + 1) + (u.M + 1) * u.L;
  hipLaunchKernelGGL((k_up_bilin<T>), dim3((n + 255) / 256, nslot), dim3(256), 0, s, u);
  return hipGetLastError();
}

template hipError_t launch_upcont<float>(const UpcontArgs<float> &, int, hipStream_t);
template hipError_t launch_upcont<double>(const UpcontArgs<double> &, int, hipStream_t);

}  // namespace tcsem
