// tcsem_kernels.hip -- gfx950 kernels of the DuFort-Frankel step: the fused E/B volume march,
// the surface level, the derived Bz volume and the receiver gather.
//
// Volume march (levels N-1 .. 1), one workgroup per tile of 64 aligned columns by BY - 2 rows
// (plus one halo row each side), one lane per (i, j) column, marching z bottom-up.  Per level it
//   (a) extends the running Bz integral of the *old* B (Bderivate.cu:227-253): Bz is never
//       stored, the E update reads it from this register sum and its LDS copy;
//   (b) prefetches B, sigma and E of the next level up into registers / the other LDS buffer;
//   (c) updates Ez (ElectricalFiled.cu:277-345) on the tile plus a +x/+y halo;
//   (d) updates Ex, Ey (ElectricalFiled.cu:89-274) (the bipole cells of a source step are
//       patched afterwards by k_src, ElectricalFiled.cu:347-475);
//   (e) updates Bx, By (Bderivate.cu:79-200) from the new E of this level and the level below.
// A level reads B_old at k and k-1 and E_new at k and k+1, so one bottom-up pass moves every
// field once: read sigma, E(3), Bx, By and write E(3), Bx, By -- 11 words per cell against the
// 13 of a stored-Bz design.  The halo ring (one column/row each side) keeps every neighbour
// read inside the workgroup's LDS; halo rows recompute Bz sums (-y side) or Ez (+y), and the two
// edge lanes fetch the columns either side of the block.
// Level 0 needs the air fields from the upward continuation of Bz(0), so the march stops at
// level 1 and writes Bz_old(0); k_surface finishes level 0 after the air planes are ready.
//
// Two-step pass (k_march2): on steady-state steps two consecutive steps share one bottom-up pass
// over levels N-1 .. 3, the second step two levels behind the first, so the state between them
// stays in registers and LDS; the top levels of both steps run through k_march and k_surface.
#include <type_traits>

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
  __device__ void save(double *ps, double *, unsigned off) const { *(double *)((char *)ps + off) = s; }
  __device__ void load(const double *ps, const double *, unsigned off) { s = *(const double *)((const char *)ps + off); }
};
template <>
struct Acc<float> {
  float s, c;
  __device__ void init() { s = 0.f; c = 0.f; }
  __device__ void save(float *ps, float *pc, unsigned off) const {
    *(float *)((char *)ps + off) = s;
    *(float *)((char *)pc + off) = c;
  }
  __device__ void load(const float *ps, const float *pc, unsigned off) {
    s = *(const float *)((const char *)ps + off);
    c = *(const float *)((const char *)pc + off);
  }
  __device__ void add(float x) {
    const float t = s + x;
    const float bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  __device__ float val() const { return s + c; }
};

// DuFort-Frankel update E <- (2l - s dt)/(2l + s dt) E + 4 dt/(2l + s dt) curl / mu0
// (ElectricalFiled.cu:138-139), written with x = s dt / (2l): E <- (1-x)/(1+x) E + c4/(1+x) curl,
// c4 = 4 dt / (2l mu0).  FP64 fields: one FP64 division per update.  FP32 fields: the reciprocal
// of 1+x from v_rcp_f32 refined by one Newton step (error below one ulp, unbiased in sign).
struct Coef64 {
  double q2, c4;  // dt/(2l), 4 dt/(2l mu0)
};
__device__ inline double df_new(double e, double sig, const Coef64 &k, double curl) {
  const double x = sig * k.q2;
  const double r = 1.0 / (1.0 + x);
  return (1.0 - x) * r * e + k.c4 * r * curl;
}
struct Coef32 {
  float q2, c4;
};
__device__ inline float df_new(float e, float sig, const Coef32 &k, float curl) {
  const float x = sig * k.q2;
  const float d = 1.0f + x;
  float r = __builtin_amdgcn_rcpf(d);
  r = r * (2.0f - d * r);
  return (1.0f - x) * r * e + k.c4 * r * curl;
}
template <typename T>
struct CoefOf;
template <>
struct CoefOf<double> {
  using type = Coef64;
};
template <>
struct CoefOf<float> {
  using type = Coef32;
};
template <typename T>
__device__ inline typename CoefOf<T>::type make_coef(double lam, double dt) {
  typename CoefOf<T>::type k;
  k.q2 = (T)(dt / (2.0 * lam));
  k.c4 = (T)(4.0 * dt / (2.0 * lam * kMu0));
  return k;
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
  bool own, need_bz, need_ez;
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

// loads/stores at a byte offset from a uniform base (global_load ... vOff, s[base])
template <typename T>
__device__ __forceinline__ T ldo(const T *__restrict__ base, unsigned off) {
  return *(const __attribute__((address_space(1))) T *)((const char *)base + off);
}
template <typename T>
__device__ __forceinline__ void sto(T *__restrict__ base, unsigned off, T v) {
  *(__attribute__((address_space(1))) T *)((char *)base + off) = v;
}

// Branch-free masked stores: a raw buffer resource over one slot's field drops any store whose
// offset lies past its size, so an invalid lane stores at kDrop instead of branching around the
// store (the host keeps every slot's field below 2 GiB).
constexpr unsigned kDrop = 0x80000000u;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9-family raw buffer: 32-bit data format
// Non-temporal cache policy (the NT bit) for the march's field stores and its single-use loads.
// The new state is read only in the next step, long after it has left L2, so streaming its stores
// keeps L2 for the current state's halo rows and columns, which neighbouring tiles read again.
// Measured against policy 0: 3-6% per step on every workload.  NT on every load instead loses
// 10-18% (it evicts the halos); only the old Ex, Ey, which no other tile reads, are loaded NT.
constexpr int kNT = 2;
__device__ inline __amdgpu_buffer_rsrc_t rsrc_of(void *base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)bytes, kRsrcWord3);
}
__device__ __forceinline__ void bsto(__amdgpu_buffer_rsrc_t r, unsigned off, bool ok, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(ok ? off : kDrop), 0, kNT);
}
__device__ __forceinline__ void bsto(__amdgpu_buffer_rsrc_t r, unsigned off, bool ok, double v) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  u32x2 w;
  w.x = (unsigned)b;
  w.y = (unsigned)(b >> 32);
  __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)(ok ? off : kDrop), 0, kNT);
}

// Loads through a raw buffer resource: an offset past the resource's size (kDrop) returns 0
// without a memory access, so a lane or wave that does not need a value drops its load without a
// branch.  AUX is the cache policy (kNT for values read once per step).
template <int AUX>
__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t r, unsigned off, float) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, AUX));
}
template <int AUX>
__device__ __forceinline__ double bld(__amdgpu_buffer_rsrc_t r, unsigned off, double) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, AUX);
  return __longlong_as_double((long long)(((unsigned long long)w.y << 32) | w.x));
}

// Per-lane geometry shared by the march and the surface kernel.  Every lane computes every
// quantity (no divergent loads); the validity masks select what is kept and stored.
template <typename T>
struct Lane {
  int i, j;
  unsigned colb;  // byte offset of the column within a level (allocation base, margin included)
  bool vEx, vEy, vEzAny, vEz, vBx, vBy, own;
  T rdx, rdx2, rdy, rdy2;
  int lxm, lxp, lym, lyp;  // clamped LDS neighbour lanes
};

template <typename T, int BY>
__device__ inline void lane_setup(const MarchArgs<T> &a, const Tile &t, Lane<T> &ln) {
  const int L = a.L, M = a.M;
  const int i = t.i, j = t.j;
  ln.i = i;
  ln.j = j;
  const int ic = min(i, L + 1), jc = min(j, M + 1);
  ln.colb = (unsigned)((jc + 1) * a.PX + (ic + a.xm)) * (unsigned)sizeof(T);
  ln.own = t.own && i <= L && j <= M;
  ln.vEx = ln.own && i < L && j >= 1 && j < M;
  ln.vEy = ln.own && i >= 1 && i < L && j < M;
  ln.vEzAny = t.need_ez && i >= 1 && i < L && j >= 1 && j < M;
  ln.vEz = ln.own && ln.vEzAny;
  ln.vBx = ln.own && i >= 1 && i < L && j < M;
  ln.vBy = ln.own && i < L && j >= 1 && j < M;
  ln.rdx = a.rdx[ic + 1];
  ln.rdx2 = a.rdx2[ic + 1];
  ln.rdy = a.rdy[jc + 1];
  ln.rdy2 = a.rdy2[jc + 1];
  ln.lxm = max(t.lx - 1, 0);
  ln.lxp = min(t.lx + 1, kMarchLanes - 1);
  ln.lym = max(t.ly - 1, 0);
  ln.lyp = min(t.ly + 1, BY - 1);
}

// One level of one column: old B, sigma and old E, and the same fields of the halo column the
// lane fetches (lane 0: column i0-1, lane 63: column i0+64, other lanes: their own column again).
template <typename T>
struct Lv {
  T bx, by, s, ex, ey, ez;
  T hbx, hby, hs, hez;
};
// The same with the conductivity as an 8-bit index into the model's table of distinct values.
template <typename T>
struct Lv8 {
  T bx, by, ex, ey, ez;
  T hbx, hby, hez;
  unsigned s, hs;
};

// Volume march, levels k_hi .. k_lo of every column (tiling and per-level steps in the header
// comment of this file).  A workgroup is one 128-byte-aligned block of 64 columns (no halo lanes
// in x) by BY rows (BY-2 owned rows and one halo row on each side).  The x-neighbours a block
// needs come from two edge lanes: lane 0 keeps a second running Bz sum for column i0-1 and
// fetches its sigma and By; lane 63 fetches column i0+64 and computes its new Ez.  They go to the
// halo slots 0 and 65 of the LDS rows.
// A model batch (uneven grids, one transmitter each) passes one MarchArgs record per slot: the
// slot's own fields, tables, sizes and tiling, with the launch's level range and the slot's step
// scalars (lam, dt, dtb, amp) from the launch's table row la.gscal.
template <typename T>
__device__ inline MarchArgs<T> slot_args(const MarchArgs<T> &la, const MarchArgs<T> &g, int member) {
  MarchArgs<T> a = g;
  const double *q = la.gscal + 4 * member;
  a.lam = q[0];
  a.dt = q[1];
  a.dtb = q[2];
  a.amp = q[3];
  a.rows = la.rows;
  a.pf = la.pf;
  a.k_lo = la.k_lo;
  a.carry = la.carry;
  if (la.carry == 1) a.k_hi = la.k_hi;  // the top launch: level 1; the lower launch: N-1 of the slot
  a.nchunk = 0;
  a.chunk_len = 0;
  return a;
}

// IB: the chunked lower launch that accumulates the chunk sums of B (a.ibw set; a.ibr may be).
// 10-row tiles (FP32) at prefetch depth 1 are compiled for five waves per SIMD: at most 96 VGPRs,
// with a few values spilled to scratch, so that two 10-wave workgroups share a CU (at the
// compiler's 108 VGPRs only one fits).  Their halo rows are 20% of the waves, against 25% for
// 8-row tiles.  The chunk-sum variant of that shape is compiled for four waves per SIMD (at most
// 128 VGPRs, no spills; at five it spilled 34): its launches, chunked because the grid offers
// few tiles, hold one workgroup per CU (measured on 256 x 256 x 160: 10 % faster).  Deeper
// 10-row prefetch keeps the compiler's register choice (about 145 VGPRs, no spills at depth 2):
// one workgroup per CU with twice the loads in flight, for grids that offer few workgroups per
// CU.  12-row tiles (FP64) are compiled for three waves per SIMD (168 VGPRs
// instead of 174-200): one 12-wave workgroup per CU with 17% halo rows, against one 8-wave
// workgroup with 25%.  10-row FP64 tiles get the same limit: 8 owned rows divide 512-row grids
// into whole rounds of one workgroup per CU (512 tiles for 512 x 512 against 416 of 12 rows).  Other shapes keep the compiler's register choice.
// S8: the conductivity comes from the 8-bit index volume a.sig8 and its table a.slut (models of
// at most 256 distinct conductivities, the table entry 0 being 0 for the zero margins): one byte
// per cell and halo column read instead of a word, the same values.
template <typename T, int BY, int PF, bool GEO, bool IB, bool S8>
__global__ void __launch_bounds__(kMarchLanes *BY)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? (BY == 10 || BY == 12 ? 3 : 1)
                                                    : (BY == 10 && PF == 1 ? (IB ? 4 : 5) : 1))))
k_march(const MarchArgs<T> la, const MarchArgs<T> *geo) {
  MarchArgs<T> a;
  int tile;
  if constexpr (GEO) {
    // model batch: a 1-D grid over the launch's work list of (member, tile) pairs, laid out so
    // that every member runs on one XCD (workgroup b goes to XCD b mod 8) and its halo rows stay
    // in that XCD's L2; padding entries have member -1
    const int2 w = la.work[blockIdx.x];
    if (w.x < 0) return;
    a = slot_args(la, geo[w.x], w.x);
    tile = w.y;
  } else {
    a = la;
    const int b = blockIdx.x;
    tile = (b & 7) * a.nper + (b >> 3);  // XCD-contiguous bands of tiles
    if (tile >= a.ntx * a.nty) return;
    if (a.gs) {  // graph replay: this step's scalars from device memory
      a.lam = a.gs[0];
      a.dt = a.gs[1];
      a.dtb = a.gs[2];
    }
  }
  extern __shared__ __align__(16) unsigned char zdyn[];  // 3 N entries of T
  constexpr int LW = kMarchLanes + 2;                   // LDS row: halo, 64 lanes, halo
  __shared__ T sBx[2][BY][LW], sBy[2][BY][LW], sS[2][BY][LW];
  __shared__ T sBz[BY][LW], sEz[BY][LW];
  const int tx = tile % a.ntx, ty = tile / a.ntx;
  // One wavefront per row (64 lanes in x), so the row index is uniform across the wave.  Telling
  // the compiler so (scalar registers for the row's tables and LDS row offsets) frees 6-15 VGPRs
  // and most spills of the FP64 and model-batch variants (measured faster there); the FP32
  // single-model variants ran 1-2% slower with it, so they keep the per-lane index.
  const int lx = threadIdx.x;
  const int ly = (GEO || sizeof(T) == 8) ? __builtin_amdgcn_readfirstlane((int)threadIdx.y) : (int)threadIdx.y;
  const int L = a.L, M = a.M, N = a.N;
  const int i0 = tx * kMarchLanes, i = i0 + lx, iR = i0 + kMarchLanes;
  const int j = ty * (BY - 2) - 1 + ly;
  const bool own = ly >= 1 && ly <= BY - 2 && i < L && j >= 0 && j < M;
  const bool vEx = own && j >= 1;
  const bool vEy = own && i >= 1;
  const bool vEzAny = ly >= 1 && i >= 1 && i < L && j >= 1 && j < M;  // tile rows and the +y halo row
  const bool vEz = own && vEzAny;
  const bool vBx = own && i >= 1;
  const bool vBy = own && j >= 1;
  const bool vEzR = ly >= 1 && iR < L && j >= 1 && j < M;  // column i0+64 (used by lane 63)
  const int ic = min(i, L + 1), jc = min(max(j, -1), M + 1);
  const int ih = lx == 0 ? i0 - 1 : (lx == kMarchLanes - 1 ? min(iR, L + 1) : ic);
  const unsigned rowe = (unsigned)((jc + 1) * a.PX + a.xm);  // element of column 0 in this row
  const unsigned colb = (rowe + ic) * (unsigned)sizeof(T);
  const unsigned hcolb = (unsigned)((int)rowe + ih) * (unsigned)sizeof(T);
  const int sl = lx + 1, slm = lx, slp = lx + 2;
  const int hsl = lx == 0 ? 0 : (lx == kMarchLanes - 1 ? LW - 1 : sl);
  const bool left = lx == 0, right = lx == kMarchLanes - 1;
  const int lym = max(ly - 1, 0), lyp = min(ly + 1, BY - 1);
  const T rdx = a.rdx[ic + 1], rdx2 = a.rdx2[ic + 1], rdy = a.rdy[jc + 1], rdy2 = a.rdy2[jc + 1];
  const T rdxL = a.rdx[max(i0 - 1, -1) + 1], rdx2R = a.rdx2[min(iR, L + 1) + 1];
  // slots in grid y: each slot's tiles run as XCD-contiguous bands one slot after another, which
  // keeps neighbouring tiles (halo rows and columns) close in time; the conductivity that the
  // slots share stays in the 256 MB last-level cache between them (measured: putting the slots of
  // a tile together in grid x made the 32-slot march 16 % slower)
  const int slot = GEO ? 0 : (int)blockIdx.y;  // a model batch's records point at slot 0 of each model
  const long long fo = (long long)slot * a.fstride - (a.PX + a.xm);  // allocation base of the slot
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
  const T *__restrict__ sig = a.sig - (a.PX + a.xm);
  const unsigned PSb = (unsigned)a.PS * (unsigned)sizeof(T);
  const unsigned fbytes = (unsigned)(a.fstride * (long long)sizeof(T));
  const __amdgpu_buffer_rsrc_t rex = rsrc_of(ex, fbytes), rey = rsrc_of(ey, fbytes), rez = rsrc_of(ez, fbytes);
  const __amdgpu_buffer_rsrc_t rbx = rsrc_of(bx, fbytes), rby = rsrc_of(by, fbytes);
  // The old Ex, Ey are used only by the tile's own rows (E updates its own cells; the halo rows
  // compute Bz sums and Ez), so the halo-row waves drop those loads, and each remaining load is
  // the value's only read in the step: non-temporal, like the stores.
  const __amdgpu_buffer_rsrc_t rexi = rsrc_of((void *)exi, fbytes), reyi = rsrc_of((void *)eyi, fbytes);
  // The -y halo row computes no Ez, so its wave drops the old-Ez loads too (that row is the tile
  // below's last owned row: a second read of it, often an L2 miss); only lane 63 uses the old Ez
  // of column i0+64.
  const __amdgpu_buffer_rsrc_t rezi = rsrc_of((void *)ezi, fbytes);
  const bool ezrow = ly >= 1;
  const bool erow = ly >= 1 && ly <= BY - 2;
  const auto cf = make_coef<T>(a.lam, a.dt);
  const T dtb = (T)a.dtb, zero = 0;

  // A ring of PF+2 register slots holds level k (cur), level k-1 (nxt) and the PF-1 levels below
  // it; each level issues the loads of level k-1-PF into the slot cur held one level earlier, so
  // a load has PF levels of work to arrive.  The level loop is unrolled by the ring length so
  // every slot is a fixed register set (no moves that would wait on loads in flight).  Levels
  // below k_lo-1 are never read (the top launch of a step owns them).
  constexpr int S = PF + 2;
  // level range: k_hi .. k_lo, or one z-chunk of it; a chunk above the bottom one starts one
  // level lower than it stores (kst) to recompute E_new there, which its first B update needs
  int khi = a.k_hi, klo = a.k_lo, kst = a.k_hi;
  const int chunk = blockIdx.z;
  if (a.nchunk > 1) {
    kst = a.k_hi - chunk * a.chunk_len;
    klo = max(kst - a.chunk_len + 1, a.k_lo);
    khi = chunk ? kst + 1 : kst;
  }
  const int lfloor = klo - 1;
  const long long po = (long long)slot * a.pstride - (a.PX + a.xm);
  using LvT = std::conditional_t<S8, Lv8<T>, Lv<T>>;
  const unsigned char *__restrict__ sig8 = S8 ? a.sig8 - (a.PX + a.xm) : nullptr;
  __shared__ T sLut[S8 ? 256 : 1];
  if constexpr (S8) {
    for (int q = ly * kMarchLanes + lx; q < 256; q += kMarchLanes * BY) sLut[q] = a.slut[q];
    __syncthreads();
  }
  constexpr unsigned kEsh = sizeof(T) == 8 ? 3 : 2;  // byte offset of an element >> kEsh = index
  auto load = [&](int lv) {
    const unsigned lb = (unsigned)max(lv, lfloor) * PSb;
    const unsigned o = lb + colb, h = lb + hcolb;
    LvT v;
    v.bx = ldo(bxi, o);
    v.by = ldo(byi, o);
    if constexpr (S8)
      v.s = ldo(sig8, o >> kEsh);
    else
      v.s = ldo(sig, o);
    v.ex = bld<kNT>(rexi, erow ? o : kDrop, T(0));
    v.ey = bld<kNT>(reyi, erow ? o : kDrop, T(0));
    v.ez = bld<0>(rezi, ezrow ? o : kDrop, T(0));
    v.hbx = ldo(bxi, h);
    v.hby = ldo(byi, h);
    if constexpr (S8)
      v.hs = ldo(sig8, h >> kEsh);
    else
      v.hs = ldo(sig, h);
    v.hez = bld<0>(rezi, ezrow && right ? h : kDrop, T(0));
    return v;
  };
  unsigned off = (unsigned)khi * PSb + colb;
  LvT R[S];
#pragma unroll
  for (int d = 0; d < S - 1; ++d) R[d] = load(khi - d);
  // dz, 1/dz, 1/(dz[k-1]+dz[k]) of every level in LDS: per-level reads then wait on the LDS
  // counter only, never on the in-order vector-memory counter that the prefetch loads share
  T *sZ = (T *)zdyn;
  for (int q = ly * kMarchLanes + lx; q < 3 * N; q += kMarchLanes * BY) sZ[q] = a.dzt[q];
  // level k_hi into the LDS buffer of its parity; the edge lanes fill the halo slots (the other
  // lanes store their own value again, so the stores need no branch)
  auto stage = [&](const int pb, const LvT &v) {
    T vs, vhs;
    if constexpr (S8) {
      vs = sLut[v.s];
      vhs = sLut[v.hs];
    } else {
      vs = v.s;
      vhs = v.hs;
    }
    sBx[pb][ly][sl] = v.bx;
    sBy[pb][ly][sl] = v.by;
    sS[pb][ly][sl] = vs;
    sBx[pb][ly][hsl] = right ? v.hbx : v.bx;
    sBy[pb][ly][hsl] = left ? v.hby : v.by;
    sS[pb][ly][hsl] = (left || right) ? vhs : vs;
  };
  stage(khi & 1, R[0]);
  Acc<T> bz, bzL;  // running Bz of the lane's column and (lane 0) of column i0-1
  bz.init();
  bzL.init();
  T ex_n1 = 0, ey_n1 = 0;  // E_new at level k+1 (zero at the bottom wall, level N)
  if (a.carry == 1) {      // resume: Bz sums of level khi+1 and E_new(khi+1) of the lower launch
    bz.load(a.cs + po, a.cc + po, colb);
    bzL.load(a.cs + po, a.cc + po, hcolb);
    ex_n1 = vEx ? ldo((const T *)ex, off + PSb) : zero;
    ey_n1 = vEy ? ldo((const T *)ey, off + PSb) : zero;
  }
  // element offsets of the lane's column and of the column i0-1 (lane 0) in a plane
  const int ecol = (int)rowe + ic, ehal = (int)rowe + ih;
  if (IB && a.ibr) {  // Bz partial sums of the chunks below from the previous step's chunk sums of B
    // every interior cell of the tile (halo rows too: the first owned row's Ex reads the Bz of the
    // row above from LDS), the cells k_zsum covers
    if (j >= 0 && j < M && i < L) {
      double sx = 0.0, sl = 0.0;
      for (int c = 0; c < chunk; ++c) {
        const double *px = a.ibr + po + (long long)(2 * c) * a.ibstride, *py = px + a.ibstride;
        sx += (px[ecol + 1] - px[ecol]) * (double)rdx + (py[ecol + a.PX] - py[ecol]) * (double)rdy;
        sl += (px[ehal + 1] - px[ehal]) * (double)rdxL + (py[ehal + a.PX] - py[ehal]) * (double)rdy;
      }
      bz.s += (T)sx;
      bzL.s += (T)sl;
    }
  } else {
    for (int c = 0; c < chunk; ++c) {  // Bz partial sums of the chunks below (k_zsum)
      bz.s += ldo((const T *)(a.zs + c * a.zstride + po), colb);
      bzL.s += ldo((const T *)(a.zs + c * a.zstride + po), hcolb);
    }
  }
  double ibx = 0.0, iby = 0.0;  // this chunk's sums of dz B (as stored) for the next step
  T bx_last = 0, by_last = 0;  // old Bx, By of level k_lo-1 after the loop
  __syncthreads();
  // Drain the set-up loads here: values loaded before the loop (the carried sums, E_new(k_hi+1))
  // would otherwise keep the loop-header state pending and cost a full wait every level.
  __builtin_amdgcn_s_waitcnt(0);

  // One level.  cur = level k, nxt = level k-1, dst = the free slot for the loads of level
  // k-1-PF.  Source cells are not applied here (a separate pass patches them), which keeps the
  // loop free of data-dependent branches, so the compiler's wait counts stay exact.
  auto level = [&](const int k, const LvT &cur, const LvT &nxt, LvT &dst) {
    const int p = k & 1, q = p ^ 1;
    const bool st = k <= kst;  // false only on the E-only level that opens an upper chunk
    const T dzk = st ? sZ[k] : zero, rdzk = sZ[N + k], rdz2k = sZ[2 * N + k];
    const T bx_k = cur.bx, by_k = cur.by, bx_m = nxt.bx, by_m = nxt.by;
    // (a) Bz_old(k) = Bz_old(k+1) + dz(k) div_h(B_old(k)), for the lane's column and column i0-1
    // (the opening level of an upper chunk starts from Bz_old(k) itself: its increment is zero)
    bz.add(dzk * ((sBx[p][ly][slp] - bx_k) * rdx + (sBy[p][lyp][sl] - by_k) * rdy));
    bzL.add(dzk * ((bx_k - cur.hbx) * rdxL + (sBy[p][lyp][0] - cur.hby) * rdy));
    const T bzk = bz.val();
    sBz[ly][sl] = bzk;
    sBz[ly][left ? 0 : sl] = left ? bzL.val() : bzk;
    // (b) level k-1 into LDS; loads of level k-1-PF
    stage(q, nxt);
    bx_last = bx_m;
    by_last = by_m;
    const T ezo = cur.ez, exo = cur.ex, eyo = cur.ey;
    dst = load(k - 1 - PF);
    __syncthreads();
    // (c) Ez(k) on the tile rows and the +y halo row; lane 63 also for column i0+64
    T ezn, ezr;
    {
      const T s4 = (T)0.25 * (sS[p][lym][slm] + sS[p][lym][sl] + sS[p][ly][slm] + sS[p][ly][sl]);
      const T curl = (by_k - sBy[p][ly][slm]) * rdx2 - (bx_k - sBx[p][lym][sl]) * rdy2;
      ezn = vEzAny ? df_new(ezo, s4, cf, curl) : zero;
    }
    {
      const T s4 = (T)0.25 * (sS[p][lym][LW - 2] + sS[p][lym][LW - 1] + sS[p][ly][LW - 2] + sS[p][ly][LW - 1]);
      const T curl = (cur.hby - by_k) * rdx2R - (cur.hbx - sBx[p][lym][LW - 1]) * rdy2;
      ezr = vEzR ? df_new(cur.hez, s4, cf, curl) : zero;
    }
    sEz[ly][sl] = ezn;
    sEz[ly][right ? LW - 1 : sl] = right ? ezr : ezn;
    bsto(rez, off, vEz && st, ezn);
    // (d) Ex(k), Ey(k)
    T exn, eyn;
    {
      const T sxk = (T)0.5 * (sS[p][lym][sl] + sS[p][ly][sl]);
      const T sxm = (T)0.5 * (sS[q][lym][sl] + sS[q][ly][sl]);
      const T ave = (T)0.5 * (sxk + sxm);
      const T curl = (bzk - sBz[lym][sl]) * rdy2 - (by_k - by_m) * rdz2k;
      exn = vEx ? df_new(exo, ave, cf, curl) : zero;
    }
    {
      const T s4 = (T)0.25 * (sS[p][ly][slm] + sS[p][ly][sl] + sS[q][ly][slm] + sS[q][ly][sl]);
      const T curl = (bx_k - bx_m) * rdz2k - (bzk - sBz[ly][slm]) * rdx2;
      eyn = vEy ? df_new(eyo, s4, cf, curl) : zero;
    }
    bsto(rex, off, vEx && st, exn);
    bsto(rey, off, vEy && st, eyn);
    __syncthreads();
    // (e) Bx(k), By(k) from E_new(k), E_new(k+1)
    const T bxn = bx_k + dtb * ((ey_n1 - eyn) * rdzk - (sEz[lyp][sl] - ezn) * rdy);
    const T byn = by_k + dtb * ((sEz[ly][slp] - ezn) * rdx - (ex_n1 - exn) * rdzk);
    bsto(rbx, off, vBx && st, bxn);
    bsto(rby, off, vBy && st, byn);
    if constexpr (IB) {  // dzk is zero on a level the launch does not store
      ibx += (double)dzk * (double)(vBx ? bxn : bx_k);
      iby += (double)dzk * (double)(vBy ? byn : by_k);
    }
    ex_n1 = exn;
    ey_n1 = eyn;
    off -= PSb;
  };
  // whole rounds of S levels (the ring returns to its starting phase), then the remainder
  int k = khi;
  for (; k - (S - 1) >= klo; k -= S) {
#pragma unroll
    for (int u = 0; u < S; ++u) level(k - u, R[u], R[(u + 1) % S], R[(u + S - 1) % S]);
  }
#pragma unroll
  for (int u = 0; u < S - 1; ++u)
    if (k - u >= klo) level(k - u, R[u], R[(u + 1) % S], R[(u + S - 1) % S]);
  if (a.carry == 2) {  // lower launch: hand the running sums of level klo to the top launch
    if (own && chunk == max(a.nchunk, 1) - 1) bz.save(a.cs + po, a.cc + po, colb);
    if (IB && own && chunk < a.nchunk - 1) {
      double *px = a.ibw + po + (long long)(2 * chunk) * a.ibstride;
      px[ecol] = ibx;
      px[a.ibstride + ecol] = iby;
    }
    return;
  }
  // Bz_old(0) for the upward continuation and the surface level (level 0 sits in buffer 0)
  bz.add(sZ[0] * ((sBx[0][ly][slp] - bx_last) * rdx + (sBy[0][lyp][sl] - by_last) * rdy));
  if (own) sto(a.bz0 + po, colb, bz.val());
}

// ---- two steps in one lower pass ----------------------------------------------------------------
// k_march2 (PairArgs): step A (old -> intermediate state) at levels N-1 .. 2 and step B
// (intermediate -> new state) at levels N-1 .. 3 in one bottom-up pass.  Step B at level m reads
// the intermediate B at m and m-1 and E at m, so it runs two levels behind step A: iteration k
// takes step A at level k and step B at level k+2, and the intermediate state below level 2 lives
// in registers and LDS only.  Per cell-update that halves the HBM traffic of the lower levels
// (read sigma and the old state once, write the new state once per two steps), for a halo ring
// two cells wide instead of one.
// Every quantity is the one-step march's expression in the same operation order, so the two
// steps give the bits of two one-step passes.  The tile is 64 lanes (columns i0-2 .. i0+61) by
// BY rows (j0-2 .. j0+BY-3): step A is exact on lanes and rows 1 .. 62 / BY-2 (the neighbours its
// curls read are one cell further out), step B on lanes 2 .. 61 and rows 2 .. BY-3, the cells a
// workgroup owns.  Cells outside the grid read and compute as zero, as memory holds them.
// Per iteration, three phases with a barrier after the first two (the third runs into the next
// iteration's first, which touches none of its LDS):
//   P1  A: Bz sum of level k (B_old of level k: y-neighbours from LDS); stage B_old and sigma of
//          level k-1; loads of level k-1-PF.   B: Bz sum of level k+2 (intermediate B).
//   P2  A: intermediate E of level k.   B: new E of level k+2 (stored).
//   P3  A: intermediate B of level k (into LDS for step B).   B: new B of level k+2 (stored).
// Step B idles on the first two iterations (levels N+1, N: dz 0, nothing stored, inputs zero)
// and step A on the last (k = 1: nothing kept), so every iteration runs the same code.  The
// lower pass never reads level 0 of the old state (the previous pair's surface kernel writes it).
// The same register of the lane to the left (l-1) or right (l+1) in the wavefront, by a DPP
// whole-wave shift (wave_shr:1, wave_shl:1); the lanes past either end read 0.  The pass's rows
// are single wavefronts, so its x-neighbours need no LDS round trip (in FP32, half its LDS reads).
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ T from_left(T v) {
  return dpp_move<kDppWaveShr1>(v);
}
template <typename T>
__device__ __forceinline__ T from_right(T v) {
  return dpp_move<kDppWaveShl1>(v);
}
template <typename T, bool S8>
struct Lv2 {
  T bx, by, ex, ey, ez;
  std::conditional_t<S8, unsigned, T> s;
};
// GEO: a model batch, one record per member (its fields, tables, sizes, intermediate state, carry
// planes and pass tiles) and a 1-D grid over the launch's work list of (member, tile) pairs; the
// launch record la carries the work list and both steps' scalar tables.
template <typename T, int BY, int PF, bool S8, bool GEO>
__global__ void __launch_bounds__(kMarchLanes *BY) k_march2(const PairArgs<T> la, const PairArgs<T> *geo) {
  PairArgs<T> pa;
  int tile;
  if constexpr (GEO) {
    const int2 w = la.a.work[blockIdx.x];
    if (w.x < 0) return;
    pa = geo[w.x];
    const double *qa = la.a.gscal + 4 * w.x, *qb = la.gscalB + 4 * w.x;
    pa.a.lam = qa[0];
    pa.a.dt = qa[1];
    pa.a.dtb = qa[2];
    pa.lamB = qb[0];
    pa.dtB = qb[1];
    pa.dtbB = qb[2];
    tile = w.y;
  } else {
    pa = la;
    const int b = blockIdx.x;
    tile = (b & 7) * pa.nper + (b >> 3);  // XCD-contiguous bands of tiles
    if (tile >= pa.ntx * pa.nty) return;
  }
  const MarchArgs<T> &a = pa.a;
  constexpr int W = kMarchLanes, TX = W - 4, TY = BY - 4;
  extern __shared__ __align__(16) unsigned char zdyn[];  // 3 N entries of T
  __shared__ T sBx[2][BY][W], sBy[2][BY][W], sS[2][BY][W];
  __shared__ T sB1x[2][BY][W], sB1y[2][BY][W];
  __shared__ T sBzA[BY][W], sEzA[BY][W], sBzB[BY][W], sEzB[BY][W];
  __shared__ T sLut[S8 ? 256 : 1];
  const int tx = tile % pa.ntx, ty = tile / pa.ntx;
  const int lx = threadIdx.x;
  const int ly = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one wave per row
  const int L = a.L, M = a.M, N = a.N;
  const int i = tx * TX - 2 + lx, j = ty * TY - 2 + ly;
  // grid cells of each component (zero outside, as memory holds them): Ex i in [0,L) j in [1,M),
  // Ey and Bx i in [1,L) j in [0,M), Ez i in [1,L) j in [1,M), By as Ex
  const bool xin = i >= 0 && i < L, xin1 = i >= 1 && i < L, yin = j >= 0 && j < M, yin1 = j >= 1 && j < M;
  const bool dEx = xin && yin1, dEy = xin1 && yin, dEz = xin1 && yin1, dBx = dEy, dBy = dEx;
  const bool own = lx >= 2 && lx < W - 2 && ly >= 2 && ly < BY - 2;
  const int ic = min(max(i, -1), L + 1), jc = min(max(j, -1), M + 1);
  const unsigned colb = (unsigned)((jc + 1) * a.PX + a.xm + ic) * (unsigned)sizeof(T);
  const int lxm = max(lx - 1, 0), lxp = min(lx + 1, W - 1), lym = max(ly - 1, 0), lyp = min(ly + 1, BY - 1);
  const T rdx = a.rdx[ic + 1], rdx2 = a.rdx2[ic + 1], rdy = a.rdy[jc + 1], rdy2 = a.rdy2[jc + 1];
  const int slot = GEO ? 0 : (int)blockIdx.y;  // a model batch's records point at each member's slot 0
  // x-neighbours by lane shifts (FP32), or from LDS (FP64: an x-neighbour from a lane shift lets
  // the compiler fuse the other product of a curl into its FMA, which would round differently
  // from the one-step march; the conductivity sums, which have no products, shift in both)
  constexpr bool kX = sizeof(T) == 4;
  const long long fo = (long long)slot * a.fstride - (a.PX + a.xm);  // allocation base of the slot
  const long long po = (long long)slot * a.pstride - (a.PX + a.xm);
  const T *__restrict__ bxi = a.bxi + fo;
  const T *__restrict__ byi = a.byi + fo;
  const T *__restrict__ sig = a.sig - (a.PX + a.xm);
  const unsigned char *__restrict__ sig8 = S8 ? a.sig8 - (a.PX + a.xm) : nullptr;
  const unsigned PSb = (unsigned)a.PS * (unsigned)sizeof(T);
  const unsigned fbytes = (unsigned)(a.fstride * (long long)sizeof(T));
  const __amdgpu_buffer_rsrc_t rex = rsrc_of(a.ex + fo, fbytes), rey = rsrc_of(a.ey + fo, fbytes),
                               rez = rsrc_of(a.ez + fo, fbytes), rbx = rsrc_of(a.bx + fo, fbytes),
                               rby = rsrc_of(a.by + fo, fbytes);
  const __amdgpu_buffer_rsrc_t mex = rsrc_of(pa.mex + fo, fbytes), mey = rsrc_of(pa.mey + fo, fbytes),
                               mez = rsrc_of(pa.mez + fo, fbytes), mbx = rsrc_of(pa.mbx + fo, fbytes),
                               mby = rsrc_of(pa.mby + fo, fbytes);
  // old E: step A computes Ex, Ey on rows 1 .. BY-2 and Ez on rows 1 .. BY-1 (row 0 only sums Bz)
  const __amdgpu_buffer_rsrc_t rexi = rsrc_of((void *)(a.exi + fo), fbytes),
                               reyi = rsrc_of((void *)(a.eyi + fo), fbytes),
                               rezi = rsrc_of((void *)(a.ezi + fo), fbytes);
  const bool erow = ly >= 1 && ly <= BY - 2, ezrow = ly >= 1;
  // the work each row's wave does (uniform branches): step A's Bz sum on rows 0 .. BY-2, its Ez on
  // rows 1 .. BY-1, its Ex, Ey and B on rows 1 .. BY-2; step B's Bz sum on rows 1 .. BY-3, its Ez
  // on rows 2 .. BY-2, its Ex, Ey and B (the stored cells) on rows 2 .. BY-3 -- what the owned
  // cells read, one ring further out for step A
  const bool rBzA = ly <= BY - 2, rA = ly >= 1 && ly <= BY - 2;
  const bool rBzB = ly >= 1 && ly <= BY - 3, rEzB = ly >= 2 && ly <= BY - 2, rB = ly >= 2 && ly <= BY - 3;
  const auto cfA = make_coef<T>(a.lam, a.dt);
  const auto cfB = make_coef<T>(pa.lamB, pa.dtB);
  const T dtbA = (T)a.dtb, dtbB = (T)pa.dtbB, zero = 0;
  constexpr int S = PF + 2;
  constexpr unsigned kEsh = sizeof(T) == 8 ? 3 : 2;
  constexpr int klo = 2, lfloor = 1;
  const int khi = N - 1;
  using LvT = Lv2<T, S8>;
  if constexpr (S8) {
    for (int q = ly * W + lx; q < 256; q += W * BY) sLut[q] = a.slut[q];
  }
  T *sZ = (T *)zdyn;
  for (int q = ly * W + lx; q < 3 * N; q += W * BY) sZ[q] = a.dzt[q];
  auto load = [&](int lv) {
    const unsigned o = (unsigned)max(lv, lfloor) * PSb + colb;
    LvT v;
    v.bx = ldo(bxi, o);
    v.by = ldo(byi, o);
    if constexpr (S8)
      v.s = ldo(sig8, o >> kEsh);
    else
      v.s = ldo(sig, o);
    v.ex = bld<0>(rexi, erow ? o : kDrop, T(0));
    v.ey = bld<0>(reyi, erow ? o : kDrop, T(0));
    v.ez = bld<0>(rezi, ezrow ? o : kDrop, T(0));
    return v;
  };
  LvT R[S];
#pragma unroll
  for (int d = 0; d < S - 1; ++d) R[d] = load(khi - d);
  auto stage = [&](const int k, const LvT &v) -> T {  // B_old and sigma of level k into LDS
    sBx[k & 1][ly][lx] = v.bx;
    sBy[k & 1][ly][lx] = v.by;
    T sv;
    if constexpr (S8)
      sv = sLut[v.s];
    else
      sv = v.s;
    sS[k & 1][ly][lx] = sv;
    return sv;  // the lane's own conductivity of level k
  };
  // step B's inputs of its idle iterations read as zero (each lane clears its own cells)
  sB1x[0][ly][lx] = sB1x[1][ly][lx] = zero;
  sB1y[0][ly][lx] = sB1y[1][ly][lx] = zero;
  __syncthreads();  // the table sLut
  T sgk = stage(khi, R[0]);  // the lane's conductivity of level k
  Acc<T> bzA, bzB;  // running Bz sums of the old and of the intermediate B
  bzA.init();
  bzB.init();
  // intermediate B of levels k+2 (c) and k+1 (n) and intermediate E of levels k+2 (c) and k+1 (n)
  // of this lane's column; new E of level k+3 (step B's level below)
  T b1xc = 0, b1yc = 0, b1xn = 0, b1yn = 0;
  T e1xc = 0, e1yc = 0, e1zc = 0, e1xn = 0, e1yn = 0, e1zn = 0;
  T e2x = 0, e2y = 0;
  // conductivity averages (Ez, Ex, Ey) of levels k+2 (c) and k+1 (n)
  T azc = 0, axc = 0, ayc = 0, azn = 0, axn = 0, ayn = 0;
  __syncthreads();
  __builtin_amdgcn_s_waitcnt(0);

  auto level = [&](const int k, const LvT &cur, const LvT &nxt, LvT &dst) {
    const int kB = k + 2, p = k & 1;  // level k+2 has the parity of k
    const bool aon = k >= klo, bon = kB <= N - 1;
    const int kBc = min(kB, N - 1);
    const T dzA = aon ? sZ[k] : zero, rdzA = sZ[N + k], rdz2A = sZ[2 * N + k];
    const T dzB = bon ? sZ[kBc] : zero, rdzB = sZ[N + kBc], rdz2B = sZ[2 * N + kBc];
    const unsigned offA = (unsigned)k * PSb + colb, offB = (unsigned)kBc * PSb + colb;
    // ---- P1
    const T bx_k = cur.bx, by_k = cur.by, bx_m = nxt.bx, by_m = nxt.by;
    T bzAk = zero, bzBk = zero;
    if (rBzA) {
      bzA.add(dzA * (((kX ? from_right(bx_k) : sBx[p][ly][lxp]) - bx_k) * rdx + (sBy[p][lyp][lx] - by_k) * rdy));
      bzAk = bzA.val();
    }
    sBzA[ly][lx] = bzAk;
    if (rBzB) {
      bzB.add(dzB * (((kX ? from_right(b1xc) : sB1x[p][ly][lxp]) - b1xc) * rdx + (sB1y[p][lyp][lx] - b1yc) * rdy));
      bzBk = bzB.val();
    }
    sBzB[ly][lx] = bzBk;
    const T sgm = stage(k - 1, nxt);  // the lane's conductivity of level k-1
    const T ezo = cur.ez, exo = cur.ex, eyo = cur.ey;
    dst = load(k - 1 - PF);
    __syncthreads();
    // ---- P2: intermediate E of level k (step A), new E of level k+2 (step B).  The conductivity
    // averages of a level are the same for both steps: step A forms them, step B takes them two
    // iterations later from registers.
    T ezA = zero, exA = zero, eyA = zero, azA = zero, axA = zero, ayA = zero;
    if (ezrow) {
      const int sp = k & 1, sq = sp ^ 1;
      const T sgkL = from_left(sgk), sgmL = from_left(sgm);
      azA = (T)0.25 * (sS[sp][lym][lxm] + sS[sp][lym][lx] + sgkL + sgk);
      {
        const T sxk = (T)0.5 * (sS[sp][lym][lx] + sgk);
        const T sxm = (T)0.5 * (sS[sq][lym][lx] + sgm);
        axA = (T)0.5 * (sxk + sxm);
      }
      ayA = (T)0.25 * (sgkL + sgk + sgmL + sgm);
      {
        const T curl = (by_k - (kX ? from_left(by_k) : sBy[p][ly][lxm])) * rdx2 - (bx_k - sBx[p][lym][lx]) * rdy2;
        const T v = df_new(ezo, azA, cfA, curl);
        ezA = dEz ? v : zero;
      }
      if (rA) {
        {
          const T curl = (bzAk - sBzA[lym][lx]) * rdy2 - (by_k - by_m) * rdz2A;
          const T v = df_new(exo, axA, cfA, curl);
          exA = dEx ? v : zero;
        }
        {
          const T curl = (bx_k - bx_m) * rdz2A - (bzAk - (kX ? from_left(bzAk) : sBzA[ly][lxm])) * rdx2;
          const T v = df_new(eyo, ayA, cfA, curl);
          eyA = dEy ? v : zero;
        }
      }
    }
    sEzA[ly][lx] = ezA;
    T ezB = zero, exB = zero, eyB = zero;
    if (rEzB) {
      {
        const T curl = (b1yc - (kX ? from_left(b1yc) : sB1y[p][ly][lxm])) * rdx2 - (b1xc - sB1x[p][lym][lx]) * rdy2;
        const T v = df_new(e1zc, azc, cfB, curl);
        ezB = bon && dEz ? v : zero;
      }
      if (rB) {
        {
          const T curl = (bzBk - sBzB[lym][lx]) * rdy2 - (b1yc - b1yn) * rdz2B;
          const T v = df_new(e1xc, axc, cfB, curl);
          exB = bon && dEx ? v : zero;
        }
        {
          const T curl = (b1xc - b1xn) * rdz2B - (bzBk - (kX ? from_left(bzBk) : sBzB[ly][lxm])) * rdx2;
          const T v = df_new(e1yc, ayc, cfB, curl);
          eyB = bon && dEy ? v : zero;
        }
      }
    }
    sEzB[ly][lx] = ezB;
    const bool stB = own && bon;
    bsto(rez, offB, stB && dEz, ezB);
    bsto(rex, offB, stB && dEx, exB);
    bsto(rey, offB, stB && dEy, eyB);
    __syncthreads();
    // ---- P3: intermediate B of level k (step A, into LDS), new B of level k+2 (step B)
    T b1xA = zero, b1yA = zero;
    if (rA) {
      const T bxA = bx_k + dtbA * ((e1yn - eyA) * rdzA - (sEzA[lyp][lx] - ezA) * rdy);
      const T byA = by_k + dtbA * (((kX ? from_right(ezA) : sEzA[ly][lxp]) - ezA) * rdx - (e1xn - exA) * rdzA);
      b1xA = dBx ? bxA : zero;
      b1yA = dBy ? byA : zero;
    }
    sB1x[p][ly][lx] = b1xA;
    sB1y[p][ly][lx] = b1yA;
    const bool stA = own && k == klo;  // level 2 of the intermediate state, for the top launches
    bsto(mex, offA, stA && dEx, exA);
    bsto(mey, offA, stA && dEy, eyA);
    bsto(mez, offA, stA && dEz, ezA);
    bsto(mbx, offA, stA && dBx, b1xA);
    bsto(mby, offA, stA && dBy, b1yA);
    if (rB) {
      const T bxB = b1xc + dtbB * ((e2y - eyB) * rdzB - (sEzB[lyp][lx] - ezB) * rdy);
      const T byB = b1yc + dtbB * (((kX ? from_right(ezB) : sEzB[ly][lxp]) - ezB) * rdx - (e2x - exB) * rdzB);
      bsto(rbx, offB, stB && dBx, bxB);
      bsto(rby, offB, stB && dBy, byB);
    }
    e2x = exB;
    e2y = eyB;
    b1xc = b1xn;
    b1yc = b1yn;
    b1xn = b1xA;
    b1yn = b1yA;
    e1xc = e1xn;
    e1yc = e1yn;
    e1zc = e1zn;
    e1xn = exA;
    e1yn = eyA;
    e1zn = ezA;
    azc = azn;
    axc = axn;
    ayc = ayn;
    azn = azA;
    axn = axA;
    ayn = ayA;
    sgk = sgm;
    // no barrier here: the next iteration's first phase writes only the Bz sums and the staging
    // buffers of level k-2, which no wave reads after this iteration's second barrier, and reads
    // the intermediate B of level k+1, which this phase does not write
  };
  int k = khi;
  for (; k - (S - 1) >= 1; k -= S) {
#pragma unroll
    for (int u = 0; u < S; ++u) level(k - u, R[u], R[(u + 1) % S], R[(u + S - 1) % S]);
  }
#pragma unroll
  for (int u = 0; u < S - 1; ++u)
    if (k - u >= 1) level(k - u, R[u], R[(u + 1) % S], R[(u + S - 1) % S]);
  // the running sums for the top launches: step A's through level 2, step B's through level 3
  if (own && xin && yin) {
    bzA.save(a.cs + po, a.cc + po, colb);
    bzB.save(pa.cs2 + po, pa.cc2 + po, colb);
  }
}

// Level 0: the air cells count as sigma = 0 and the level above is BxAir/ByAir at -dz0/2
// (ElectricalFiled.cu:107-139, 204-234).
template <typename T, int BY, bool GEO>
__global__ void __launch_bounds__(kMarchLanes *BY) k_surface(const MarchArgs<T> la, const MarchArgs<T> *geo) {
  MarchArgs<T> a;
  if constexpr (GEO) {
    a = slot_args(la, geo[blockIdx.y], (int)blockIdx.y);
  } else {
    a = la;
    if (a.gs) {  // graph replay: this step's scalars from device memory
      a.lam = a.gs[0];
      a.dt = a.gs[1];
      a.dtb = a.gs[2];
      a.amp = a.gs[3];
    }
  }
  __shared__ T sBx[BY][kMarchLanes], sBy[BY][kMarchLanes], sS[BY][kMarchLanes];
  __shared__ T sBz[BY][kMarchLanes], sEz[BY][kMarchLanes];
  Tile t;
  if ((GEO && (int)(blockIdx.x >> 3) >= a.sper) || !tile_of<BY>(a.stx, a.sty, a.sper, t)) return;
  Lane<T> ln;
  lane_setup<T, BY>(a, t, ln);
  const int lx = t.lx, ly = t.ly, i = ln.i, j = ln.j;
  const int L = a.L, M = a.M;
  const int slot = GEO ? 0 : (int)blockIdx.y;
  const long long fo = (long long)slot * a.fstride - (a.PX + a.xm);
  const long long po = (long long)slot * a.pstride - (a.PX + a.xm);
  const unsigned off = ln.colb, PSb = (unsigned)a.PS * (unsigned)sizeof(T);
  T *__restrict__ ex = a.ex + fo;
  T *__restrict__ ey = a.ey + fo;
  T *__restrict__ ez = a.ez + fo;
  T *__restrict__ bx = a.bx + fo;
  T *__restrict__ by = a.by + fo;
  const T bx0 = ldo(a.bxi + fo, off), by0 = ldo(a.byi + fo, off), s0 = ldo(a.sig - (a.PX + a.xm), off);
  const bool bzv = t.need_bz && i >= 0 && i < L && j >= 0 && j < M;
  const T bzc = bzv ? ldo(a.bz0 + po, off) : (T)0;
  const T exo = ldo(a.exi + fo, off), eyo = ldo(a.eyi + fo, off), ezo = ldo(a.ezi + fo, off);
  const T ex1 = ldo((const T *)ex, off + PSb), ey1 = ldo((const T *)ey, off + PSb);  // E_new(1)
  const T bya = ldo(a.byair + po, off), bxa = ldo(a.bxair + po, off);
  sBx[ly][lx] = bx0;
  sBy[ly][lx] = by0;
  sS[ly][lx] = s0;
  sBz[ly][lx] = bzc;
  const int sfirst = a.src_first ? a.src_first[slot] : 0, scount = a.src_count ? a.src_count[slot] : 0;
  const auto cf = make_coef<T>(a.lam, a.dt);
  const T dtb = (T)a.dtb, rdz0 = a.rdz[0], rdz20 = a.rdz2[0], zero = 0;
  const T rdx = ln.rdx, rdx2 = ln.rdx2, rdy = ln.rdy, rdy2 = ln.rdy2;
  const int lxm = ln.lxm, lxp = ln.lxp, lym = ln.lym, lyp = ln.lyp;
  __syncthreads();
  T ezn;
  {
    const T s4 = (T)0.25 * (sS[lym][lxm] + sS[lym][lx] + sS[ly][lxm] + sS[ly][lx]);
    const T curl = (by0 - sBy[ly][lxm]) * rdx2 - (bx0 - sBx[lym][lx]) * rdy2;
    ezn = apply_src(df_new(ezo, s4, cf, curl), 2, 0, j, i, a, sfirst, scount);
    ezn = ln.vEzAny ? ezn : zero;
  }
  sEz[ly][lx] = ezn;
  if (ln.vEz) sto(ez, off, ezn);
  T exn, eyn;
  {
    const T sxk = (T)0.5 * (sS[lym][lx] + sS[ly][lx]);
    const T ave = (T)0.5 * (sxk + zero);
    const T curl = (bzc - sBz[lym][lx]) * rdy2 - (by0 - bya) * rdz20;
    exn = apply_src(df_new(exo, ave, cf, curl), 0, 0, j, i, a, sfirst, scount);
    exn = ln.vEx ? exn : zero;
  }
  {
    const T s4 = (T)0.25 * (sS[ly][lxm] + sS[ly][lx]);
    const T curl = (bx0 - bxa) * rdz20 - (bzc - sBz[ly][lxm]) * rdx2;
    eyn = apply_src(df_new(eyo, s4, cf, curl), 1, 0, j, i, a, sfirst, scount);
    eyn = ln.vEy ? eyn : zero;
  }
  if (ln.vEx) sto(ex, off, exn);
  if (ln.vEy) sto(ey, off, eyn);
  __syncthreads();
  const T bxn = bx0 + dtb * ((ey1 - eyn) * rdz0 - (sEz[lyp][lx] - ezn) * rdy);
  const T byn = by0 + dtb * ((sEz[ly][lxp] - ezn) * rdx - (ex1 - exn) * rdz0);
  if (ln.vBx) sto(bx, off, bxn);
  if (ln.vBy) sto(by, off, byn);
}

// Source cells of a source step at levels 1 .. N-1, after the volume march (level 0 stays with
// k_surface).  First the E cells take E - amp w (ElectricalFiled.cu:347-475; the weights of
// coincident table entries are summed in table order, the order in which the reference's per-cell
// loop adds them), then every Bx/By cell whose curl reads one of them is recomputed from B_old and
// the patched E with the march's own expression (Bderivate.cu:79-200).  A chunked lower launch has
// already added dz(k) B(k) of these cells to its chunk sums (a.ibw set): each recomputed cell moves
// its chunk's sum by dz(k) (B_new - B_march), so that the next step may still start from the sums.
// One workgroup per slot (GEO: per model-batch member).  The host lists each cell once and the B
// cells of one (component, column) as one run, levels ascending, which one thread walks: no two
// threads touch one cell or one chunk sum.
template <typename T, bool GEO>
__global__ void __launch_bounds__(256) k_src(const MarchArgs<T> la, const MarchArgs<T> *geo) {
  MarchArgs<T> a;
  int s;
  if constexpr (GEO) {
    a = slot_args(la, geo[blockIdx.x], (int)blockIdx.x);
    s = 0;  // a member's lists are those of its slot 0
  } else {
    a = la;
    s = (int)blockIdx.x;
  }
  if (!a.fixe || a.amp == 0.0) return;
  const int n = a.fixn;
  const long long PS = a.PS, PX = a.PX;
  const long long fo = GEO ? 0LL : (long long)s * a.fstride, po = GEO ? 0LL : (long long)s * a.pstride;
  T *const f[3] = {a.ex + fo, a.ey + fo, a.ez + fo};
  const int e0 = a.fixidx[s], ne = a.fixidx[n + s];
  for (int q = threadIdx.x; q < ne; q += blockDim.x) {
    const SrcCell sc = a.fixe[e0 + q];
    T *e = f[sc.comp] + (long long)sc.k * PS + (long long)sc.j * PX + sc.i;
    *e = (T)((double)*e - a.amp * sc.w);
  }
  __syncthreads();  // workgroup-scope release/acquire: the patched E is visible to every thread
  const T dtb = (T)a.dtb;
  const T *ex = f[0], *ey = f[1], *ez = f[2];
  const T *bxi = a.bxi + fo, *byi = a.byi + fo;
  T *bx = a.bx + fo, *by = a.by + fo;
  const int g0 = a.fixidx[2 * n + s], ng = a.fixidx[3 * n + s];
  for (int g = threadIdx.x; g < ng; g += blockDim.x) {
    const int2 run = a.fixg[g0 + g];
    for (int q = run.x; q < run.x + run.y; ++q) {
      const SrcCell sc = a.fixb[q];
      const long long o = (long long)sc.k * PS + (long long)sc.j * PX + sc.i;
      const T rdzk = a.rdz[sc.k];
      T old, nw;
      if (sc.comp == 3) {
        old = bx[o];
        nw = bxi[o] + dtb * ((ey[o + PS] - ey[o]) * rdzk - (ez[o + PX] - ez[o]) * a.rdy[sc.j + 1]);
        bx[o] = nw;
      } else {
        old = by[o];
        nw = byi[o] + dtb * ((ez[o + 1] - ez[o]) * a.rdx[sc.i + 1] - (ex[o + PS] - ex[o]) * rdzk);
        by[o] = nw;
      }
      if (a.ibw && sc.k >= 2) {  // levels 2 .. N-1 belong to the lower launch's chunks
        const int c = (a.k_hi - sc.k) / a.chunk_len;
        if (c < a.nchunk - 1) {
          double *p = a.ibw + po + (long long)(2 * c + (sc.comp == 4)) * a.ibstride + (long long)sc.j * PX + sc.i;
          *p += (double)a.dzt[sc.k] * ((double)nw - (double)old);
        }
      }
    }
  }
}

// Bz partial sums of the z-chunks of a split lower march: for chunk c (levels k_hi - c*chunk_len
// down to the next chunk) and every column, the sum over the chunk of dz(k) div_h(B_old(k)) in
// the order the march adds it (Bderivate.cu:227-253).  Chunk c of the march starts from the sum
// of the partials of chunks 0 .. c-1.  Grid: x over columns, y rows, z = chunk + (nchunk-1) slot.
template <typename T>
__global__ void __launch_bounds__(256) k_zsum(const MarchArgs<T> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  const int c = blockIdx.z % (a.nchunk - 1), slot = blockIdx.z / (a.nchunk - 1);
  if (i >= a.L) return;
  const long long fo = (long long)slot * a.fstride + (long long)j * a.PX + i;
  const T *bx = a.bxi + fo, *by = a.byi + fo;
  const T rdx = a.rdx[i + 1], rdy = a.rdy[j + 1];
  const int hi = a.k_hi - c * a.chunk_len, lo = hi - a.chunk_len + 1;
  Acc<T> acc;
  acc.init();
#pragma unroll 4
  for (int k = hi; k >= lo; --k) {
    const long long o = (long long)k * a.PS;
    acc.add(a.dzt[k] * ((bx[o + 1] - bx[o]) * rdx + (by[o + a.PX] - by[o]) * rdy));
  }
  a.zs[c * a.zstride + (long long)slot * a.pstride + (long long)j * a.PX + i] = acc.val();
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

// One workgroup (kGatherThreads lanes) per receiver and component of the group.  The 4 x 4 x 4
// stencil values are loaded by 64 lanes at once; for Bz (never stored) each of the 16 stencil
// columns gets 16 lanes that split the bottom-up sum of dz div_h(Bx, By) below the stencil, and one
// lane per column then adds the four stencil levels.  The 16 x-, 4 y- and 1 z-splines run on 16,
// 4 and 1 lanes.
constexpr int kGatherThreads = 256;
template <typename T>
__global__ void __launch_bounds__(kGatherThreads) k_gather(const GatherArgs g, const MarchArgs<T> a) {
  __shared__ double sv[64];        // stencil values [ik][ij][q]
  __shared__ double spart[256];    // Bz partial sums [column][part]
  __shared__ double sy1[16], sy2[4];
  const int slot = blockIdx.y;
  const int nr = g.rx_count[slot];
  const int r = blockIdx.x / g.ncomp_group, cg = blockIdx.x % g.ncomp_group;
  if (r >= nr) return;  // uniform over the workgroup
  const int lane = threadIdx.x;
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
  const long long fo = (long long)slot * g.fstride;
  if (comp < 5) {
    if (lane < 64) {
      const int ik = lane >> 4, ij = (lane >> 2) & 3, q = lane & 3;
      const T *f = (const T *)g.f[comp] + fo;
      sv[lane] = (double)f[(long long)(ik + s3) * g.PS + (long long)(ij + s2) * g.PX + s1 + q];
    }
  } else {
    // Bz at levels s3..s3+3 of the 4 x 4 columns: bottom-up running sums of div_h(Bx, By)
    const T *bxf = (const T *)g.f[3] + fo, *byf = (const T *)g.f[4] + fo;
    const int col = lane >> 4, part = lane & 15, ij = col >> 2, q = col & 3;
    const int ii = s1 + q, jj = s2 + ij;
    const T rdx = a.rdx[ii + 1], rdy = a.rdy[jj + 1];
    const long long base = (long long)jj * g.PX + ii;
    auto div = [&](int k) {
      const long long o = (long long)k * g.PS + base;
      return (T)(a.dzt[k] * ((bxf[o + 1] - bxf[o]) * rdx + (byf[o + g.PX] - byf[o]) * rdy));
    };
    Acc<T> acc;
    acc.init();
    for (int k = g.N - 1 - part; k >= s3 + 4; k -= 16) acc.add(div(k));
    spart[lane] = (double)acc.val();
    __syncthreads();
    if (part == 0) {
      double run = 0.0;
      for (int p = 15; p >= 0; --p) run += spart[col * 16 + p];
      for (int ik = 3; ik >= 0; --ik) {
        const int k = s3 + ik;
        if (k < g.N) run += (double)div(k);
        sv[ik * 16 + ij * 4 + q] = k < g.N ? run : 0.0;
      }
    }
  }
  __syncthreads();
  double tx[4], ty[4], tz[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    tx[q] = xe ? X[s1 + q] - 0.5 * dx[s1 + q] : X[s1 + q];
    ty[q] = ye ? Y[s2 + q] - 0.5 * dy[s2 + q] : Y[s2 + q];
    tz[q] = ze ? Z[s3 + q] - 0.5 * dz[s3 + q] : Z[s3 + q];
  }
  if (lane < 16) sy1[lane] = spline4(tx, &sv[lane * 4], x);  // lane = ik*4 + ij
  __syncthreads();
  if (lane < 4) sy2[lane] = spline4(ty, &sy1[lane * 4], y);
  __syncthreads();
  if (lane != 0) return;
  const double v = spline4(tz, sy2, z);
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
// Raises a march kernel's dynamic-LDS limit once per size (static + dynamic may pass 64 KiB).
template <typename T, int BY, int PF, bool GEO, bool IB, bool S8>
hipError_t march_lds(size_t bytes) {
  static size_t allowed = 0;  // per instantiation; only grows
  if (bytes <= allowed) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute((const void *)k_march<T, BY, PF, GEO, IB, S8>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) allowed = bytes;
  return e;
}
template <typename T, int BY, int PF, bool GEO, bool IB, bool S8 = false>
hipError_t launch_march_v(const MarchArgs<T> &a, const MarchArgs<T> *geo, dim3 grd, dim3 blk, hipStream_t s,
                          hipEvent_t e0, hipEvent_t e1) {
  const size_t zl = (size_t)3 * a.N * sizeof(T);
  hipError_t e;
  if ((e = march_lds<T, BY, PF, GEO, IB, S8>(zl)) != hipSuccess) return e;
  launch_k(k_march<T, BY, PF, GEO, IB, S8>, grd, blk, (std::uint32_t)zl, s, e0, e1, a, geo);
  return hipGetLastError();
}
// geo == nullptr: one model (slots in grid y share a's geometry); else a model batch (a.N, a.nper:
// the largest over the slots' records, which carry the rest).  A chunked launch with a.ibw set
// runs the variant that accumulates the chunk sums of B; with a.sig8 set (a model batch: when
// every member's record carries its index), the variant that reads the 8-bit conductivity index.
template <typename T, int BY, int PF>
hipError_t launch_march_pf(const MarchArgs<T> &a, const MarchArgs<T> *geo, dim3 grd, dim3 blk, hipStream_t s,
                           hipEvent_t e0, hipEvent_t e1) {
  if (geo)
    return a.sig8 ? launch_march_v<T, BY, PF, true, false, true>(a, geo, grd, blk, s, e0, e1)
                  : launch_march_v<T, BY, PF, true, false>(a, geo, grd, blk, s, e0, e1);
  const bool ib = a.ibw && a.nchunk > 1;
  if (a.sig8)
    return ib ? launch_march_v<T, BY, PF, false, true, true>(a, geo, grd, blk, s, e0, e1)
              : launch_march_v<T, BY, PF, false, false, true>(a, geo, grd, blk, s, e0, e1);
  return ib ? launch_march_v<T, BY, PF, false, true>(a, geo, grd, blk, s, e0, e1)
            : launch_march_v<T, BY, PF, false, false>(a, geo, grd, blk, s, e0, e1);
}
template <typename T, int BY>
hipError_t launch_march_by(const MarchArgs<T> &a, const MarchArgs<T> *geo, int nslot, hipStream_t s, hipEvent_t e0,
                           hipEvent_t e1) {
  dim3 blk(kMarchLanes, BY), grd(8 * a.nper, nslot, max(a.nchunk, 1));
  switch (a.pf) {
    case 1: return launch_march_pf<T, BY, 1>(a, geo, grd, blk, s, e0, e1);
    case 2: return launch_march_pf<T, BY, 2>(a, geo, grd, blk, s, e0, e1);
    case 3: return launch_march_pf<T, BY, 3>(a, geo, grd, blk, s, e0, e1);
    case 4: return launch_march_pf<T, BY, 4>(a, geo, grd, blk, s, e0, e1);
  }
  return hipErrorInvalidValue;
}
template <typename T>
hipError_t launch_march(const MarchArgs<T> &a, int nslot, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                        const MarchArgs<T> *geo) {
  switch (a.rows) {
    case 4: return launch_march_by<T, 4>(a, geo, nslot, s, e0, e1);
    case 8: return launch_march_by<T, 8>(a, geo, nslot, s, e0, e1);
    case 10: return launch_march_by<T, 10>(a, geo, nslot, s, e0, e1);
    case 12:  // FP64 only
      if constexpr (sizeof(T) == 8) return launch_march_by<T, 12>(a, geo, nslot, s, e0, e1);
      break;
    case 16:  // FP32 only (FP64 would not fit its registers)
      if constexpr (sizeof(T) == 4) return launch_march_by<T, 16>(a, geo, nslot, s, e0, e1);
      break;
  }
  return hipErrorInvalidValue;
}
// 8, 10, 12, 14 and 16 rows (FP32), 8, 10, 11 and 12 (FP64: more rows would spill its registers)
int pair_rows_supported(int prec, int rows) {
  if (rows == 8 || rows == 10 || rows == 12) return rows;
  if (prec == 32) return rows == 14 || rows == 16 ? rows : 0;
  return rows == 11 ? rows : 0;
}
template <typename T, int BY, int PF, bool S8, bool GEO>
hipError_t launch_march2_v(const PairArgs<T> &p, int nslot, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                           const PairArgs<T> *geo) {
  const size_t zl = (size_t)3 * p.a.N * sizeof(T);
  static size_t allowed = 0;  // per instantiation; only grows
  if (zl > allowed) {
    const hipError_t e = hipFuncSetAttribute((const void *)k_march2<T, BY, PF, S8, GEO>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)zl);
    if (e != hipSuccess) return e;
    allowed = zl;
  }
  dim3 blk(kMarchLanes, BY), grd(8 * p.nper, GEO ? 1 : nslot);
  launch_k(k_march2<T, BY, PF, S8, GEO>, grd, blk, (std::uint32_t)zl, s, e0, e1, p, geo);
  return hipGetLastError();
}
template <typename T, int BY>
hipError_t launch_march2_by(const PairArgs<T> &p, int nslot, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                            const PairArgs<T> *geo) {
  const bool s8 = p.a.sig8 != nullptr;
  if (geo) {  // model batches run prefetch depth 1
    if (p.a.pf != 1) return hipErrorInvalidValue;
    return s8 ? launch_march2_v<T, BY, 1, true, true>(p, nslot, s, e0, e1, geo)
              : launch_march2_v<T, BY, 1, false, true>(p, nslot, s, e0, e1, geo);
  }
  switch (p.a.pf) {
    case 1:
      return s8 ? launch_march2_v<T, BY, 1, true, false>(p, nslot, s, e0, e1, geo)
                : launch_march2_v<T, BY, 1, false, false>(p, nslot, s, e0, e1, geo);
    case 2:
      return s8 ? launch_march2_v<T, BY, 2, true, false>(p, nslot, s, e0, e1, geo)
                : launch_march2_v<T, BY, 2, false, false>(p, nslot, s, e0, e1, geo);
  }
  return hipErrorInvalidValue;
}
// geo == nullptr: one model (slots in grid y); else a model batch (p: the work list, the scalar
// tables of both steps, the largest N; geo: one record per member).
template <typename T>
hipError_t launch_march2(const PairArgs<T> &p, int nslot, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                         const PairArgs<T> *geo) {
  if (p.a.N < 6 || (!geo && (p.ntx < 1 || p.nty < 1)) || (geo && (!p.a.work || !p.gscalB)))
    return hipErrorInvalidValue;
  switch (p.rows) {
    case 8: return launch_march2_by<T, 8>(p, nslot, s, e0, e1, geo);
    case 10: return launch_march2_by<T, 10>(p, nslot, s, e0, e1, geo);
    case 11:
      if constexpr (sizeof(T) == 8) return launch_march2_by<T, 11>(p, nslot, s, e0, e1, geo);
      break;
    case 12: return launch_march2_by<T, 12>(p, nslot, s, e0, e1, geo);
    case 14:
      if constexpr (sizeof(T) == 4) return launch_march2_by<T, 14>(p, nslot, s, e0, e1, geo);
      break;
    case 16:
      if constexpr (sizeof(T) == 4) return launch_march2_by<T, 16>(p, nslot, s, e0, e1, geo);
      break;
  }
  return hipErrorInvalidValue;
}
template <typename T>
hipError_t launch_zsum(const MarchArgs<T> &a, int nslot, hipStream_t s, hipEvent_t e0) {
  launch_k(k_zsum<T>, dim3((a.L + 255) / 256, a.M, (a.nchunk - 1) * nslot), dim3(256), 0u, s, e0, nullptr, a);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_surface(const MarchArgs<T> &a, int nslot, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                          const MarchArgs<T> *geo) {
  dim3 blk(kMarchLanes, kSurfaceRows), grd(8 * a.sper, nslot);
  if (geo)
    launch_k(k_surface<T, kSurfaceRows, true>, grd, blk, 0u, s, e0, e1, a, geo);
  else
    launch_k(k_surface<T, kSurfaceRows, false>, grd, blk, 0u, s, e0, e1, a, geo);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_src(const MarchArgs<T> &a, int nslot, hipStream_t s, const MarchArgs<T> *geo) {
  if (geo)
    hipLaunchKernelGGL((k_src<T, true>), dim3(nslot), dim3(256), 0, s, a, geo);
  else
    hipLaunchKernelGGL((k_src<T, false>), dim3(nslot), dim3(256), 0, s, a, geo);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_bz_volume(const MarchArgs<T> &a, T *bz, int slot, hipStream_t s) {
  dim3 blk(64), grd((a.L + 63) / 64, a.M);
  hipLaunchKernelGGL((k_bz_volume<T>), grd, blk, 0, s, a, bz, slot);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_gather(const GatherArgs &g, int is_b, int nslot, const MarchArgs<T> &a, hipStream_t s,
                         hipEvent_t e0, hipEvent_t e1) {
  (void)is_b;
  dim3 blk(kGatherThreads), grd(g.nrmax * g.ncomp_group, nslot);
  launch_k(k_gather<T>, grd, blk, 0u, s, e0, e1, g, a);
  return hipGetLastError();
}
__global__ void k_put_scalars(double *dst, const GraphScalars s, int n) {
  const int t = threadIdx.x;
  if (t < n) dst[t] = s.v[t];
}
hipError_t launch_put_scalars(double *dst, const GraphScalars &s, int nsteps, hipStream_t st) {
  if (nsteps < 1 || nsteps > kGraphMaxSteps) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_put_scalars, dim3(1), dim3(4 * kGraphMaxSteps), 0, st, dst, s, 4 * nsteps);
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

template hipError_t launch_march<float>(const MarchArgs<float> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                        const MarchArgs<float> *);
template hipError_t launch_march<double>(const MarchArgs<double> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                         const MarchArgs<double> *);
template hipError_t launch_march2<float>(const PairArgs<float> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                         const PairArgs<float> *);
template hipError_t launch_march2<double>(const PairArgs<double> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                          const PairArgs<double> *);
template hipError_t launch_surface<float>(const MarchArgs<float> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                          const MarchArgs<float> *);
template hipError_t launch_surface<double>(const MarchArgs<double> &, int, hipStream_t, hipEvent_t, hipEvent_t,
                                           const MarchArgs<double> *);
template hipError_t launch_src<float>(const MarchArgs<float> &, int, hipStream_t, const MarchArgs<float> *);
template hipError_t launch_src<double>(const MarchArgs<double> &, int, hipStream_t, const MarchArgs<double> *);
template hipError_t launch_zsum<float>(const MarchArgs<float> &, int, hipStream_t, hipEvent_t);
template hipError_t launch_zsum<double>(const MarchArgs<double> &, int, hipStream_t, hipEvent_t);
template hipError_t launch_bz_volume<float>(const MarchArgs<float> &, float *, int, hipStream_t);
template hipError_t launch_bz_volume<double>(const MarchArgs<double> &, double *, int, hipStream_t);
template hipError_t launch_gather<float>(const GatherArgs &, int, int, const MarchArgs<float> &, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_gather<double>(const GatherArgs &, int, int, const MarchArgs<double> &, hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace tcsem
