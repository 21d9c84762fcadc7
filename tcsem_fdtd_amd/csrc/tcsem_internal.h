// tcsem_internal.h -- types shared by the host runtime (tcsem_host.cpp) and the gfx950 kernels.
//
// Field layout in HBM (per field, per batch slot): a zero margin around a pitched (k, j, i)
// volume, x fastest.  Element (k, j, i) lives at origin[k*PS + j*PX + i] with
// origin = allocation + PX + XM, XM = 128 bytes of elements, so column 0 of every row starts a
// 128-byte line (whole-line stores from the march) and the halo indices i = -1 and j = -1 are
// addressable and read as zero.  PX >= L + 2 + XM (rounded to 64 elements), PY = M + 3 rows per
// level, PS = PX * PY; every field has N + 1 levels.  The conductivity uses the same layout (N
// levels, shared by all slots).  The surface planes (Bz0, BxAir, ByAir) use one level of the same
// pitch.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tcsem {

constexpr double kPi = 3.141592653589793238;  // FTDT.h:39
constexpr double kMu0 = 4 * kPi * 1.e-7;      // main.cu:55
constexpr int kNPad = 4;                      // FTDT.h:40

constexpr int kLineBytes = 128;               // x margin of the layout: one cache line

// Volume march tile: 64 aligned columns (no halo lanes in x; two edge lanes fetch the
// neighbouring columns) by BY rows in y (BY - 2 owned rows + one halo row each side); BY is chosen
// per grid and batch size.  The surface kernel tiles 64 lanes as 62 owned columns + one halo
// column each side by kSurfaceRows rows.
constexpr int kMarchLanes = 64;
constexpr int kSurfaceOwnX = kMarchLanes - 2;
constexpr int kSurfaceRows = 8;
constexpr int kMaxChunks = 16;  // z-chunks of one lower march launch

// One injected E cell: E[comp](k, j, i) -= amp * w each source step (ElectricalFiled.cu:347-475).
// The same record lists the B cells a source step recomputes (comp 3 Bx, 4 By; w unused).
struct SrcCell {
  int comp;  // 0 Ex, 1 Ey, 2 Ez (3 Bx, 4 By in the recompute list)
  int k, j, i;
  double w;
};

// The field state is double-buffered: a step reads the old state (the *i pointers) and writes
// the new one (the *o pointers), so workgroups never see a neighbour's half-updated halo.
template <typename T>
struct MarchArgs {
  const T *exi, *eyi, *ezi, *bxi, *byi;  // old state, origin of slot 0
  T *ex, *ey, *ez, *bx, *by;             // new state, origin of slot 0
  const T *sig;               // origin, shared by all slots
  // the same conductivity as an 8-bit index (origin, same layout, one byte per cell) into the
  // table slut of the model's distinct values (entry 0 is 0): null when the model has more
  // than 255 distinct nonzero conductivities (or the index is switched off)
  const unsigned char *sig8;
  const T *slut;
  T *bz0;                     // Bz at level 0 (plane origin, slot 0)
  const T *bxair, *byair;     // air planes (origin, slot 0)
  long long fstride;          // elements between slots of one field
  long long pstride;          // elements between slots of one plane
  long long PS;               // plane stride of a volume
  int L, M, N, PX;
  int xm;                     // x margin in elements (origin = allocation + PX + xm)
  int rows;                   // workgroup rows of the march (tile height + 2): 4, 8 or 16
  int pf;                     // march prefetch depth in levels: 1 .. 4
  int ntx, nty, nper;         // march tiles in x, y; tiles per XCD group
  int stx, sty, sper;         // the same for the surface kernel
  const T *rdx, *rdx2, *rdy, *rdy2;  // 1/dx[i], 1/(dx[i-1]+dx[i]), same for y (index+1 origin)
  const T *dzt, *rdz, *rdz2;         // dz[k], 1/dz[k], 1/(dz[k-1]+dz[k]) (rdz2[0] = 1/(2 dz[0]))
  double lam, dt, dtb, amp;          // step scalars (FP64, main.cu:223-260)
  const SrcCell *src;                // all cells of all slots
  const int *src_first, *src_count;  // per slot
  int src_kmin, src_kmax;            // level range holding any source cell
  // Level range of one march launch (k_hi down to k_lo).  A step may run as a lower launch
  // (N-1 .. 2) that stores the running Bz sums (carry 2) and a top launch (level 1) that resumes
  // from them (carry 1) and writes Bz_old(0); carry 0 marches N-1 .. 1 in one launch.
  int k_hi, k_lo, carry;
  T *cs, *cc;                        // carry planes (sum, FP32 compensation), plane origin, slot 0
  // The lower launch may split its levels into nchunk z-chunks of chunk_len levels (grid z):
  // chunk c stores levels k_hi - c*chunk_len down, starting from the Bz partial sums of the
  // chunks below it (k_zsum, planes zs + c*zstride, plane origin, slot 0).
  int nchunk, chunk_len;
  T *zs;
  long long zstride;
  // model batch, launch record only: lam, dt, dtb, amp of each member for this step (4 FP64 per
  // member); the march's work list of (member, tile) per workgroup (member -1: none)
  const double *gscal;
  const int2 *work;
  // graph replay: lam, dt, dtb, amp of this step from device memory (4 FP64) instead of the
  // scalars above, which a captured launch would freeze (null: use the scalars)
  const double *gs;
  // Chunk sums of B (chunked lower launches): with ibw set, chunk c < nchunk-1 stores, per owned
  // column, the FP64 sums over its stored levels of dz(k) Bx(k) and dz(k) By(k) as the step leaves
  // them in memory (planes ibw + (2c) ibstride and + (2c+1) ibstride, plane origin, slot 0).  With
  // ibr set (the previous step's sums, same chunking), chunk c starts from the divergence of the
  // sums of chunks 0 .. c-1 instead of the k_zsum partials: sum_k dz(k) div_h(B(k)) =
  // div_h(sum_k dz(k) B(k)), as dz is uniform over a level.
  const double *ibr;
  double *ibw;
  long long ibstride;
  // Source pass of a source step (levels 1 .. N-1, k_src): per slot the merged E cells (each cell
  // once, the weights of coincident table entries summed) fixe[fixidx[s] ..] (count
  // fixidx[fixn + s]), and the Bx/By cells whose curl reads one of them, grouped by column:
  // fixg[fixidx[2 fixn + s] ..] (count fixidx[3 fixn + s]) holds (first, count) runs of fixb, one
  // run per (component, j, i), levels ascending.  Null fixe: no source pass.
  const SrcCell *fixe, *fixb;
  const int2 *fixg;
  const int *fixidx;
  int fixn;
};

// Two steps in one lower pass (k_march2): step A advances the old state (a.exi ..) to the
// intermediate state, step B the intermediate state to the new one (a.ex ..), two levels behind
// step A in the same bottom-up pass, so the intermediate state never goes through HBM below level
// 2.  Step A stores its level 2 into the intermediate buffer (mex ..) and its running Bz sums
// through level 2 into a.cs/a.cc; step B stores levels N-1 .. 3 and its sums through level 3 into
// cs2/cc2.  The top levels of both steps are the ordinary top launches and surface kernels.
// Tiles: 60 owned columns (64 lanes: two halo columns each side) by rows - 4 owned rows (two halo
// rows each side); ntx, nty, nper count them (tiles per XCD group).
template <typename T>
struct PairArgs {
  MarchArgs<T> a;                  // step A's scalars; old and new state; carry planes of step A
  T *mex, *mey, *mez, *mbx, *mby;  // intermediate state (origin, slot 0)
  T *cs2, *cc2;                    // carry planes of step B (plane origin, slot 0)
  double lamB, dtB, dtbB;          // step B's scalars
  int rows, ntx, nty, nper;
  const double *gscalB;            // model batch, launch record only: step B's row of the scalar table
};

// Receiver gather descriptor (EFInterp2Rx.cu:60-470) for one field group (E or B) of one event.
struct GatherArgs {
  const void *f[6];          // Ex Ey Ez Bx By (origin, slot 0); Bz entry unused
  long long fstride, PS;
  int L, M, N, PX;
  const double *X, *Y, *Z, *dx, *dy, *dz;  // host-equal FP64 grid
  const int *rx_local;       // per slot: nr local receiver indices into the receiver tables
  const int *rx_first;       // per slot offset into rx_local
  const int *rx_count;       // per slot count
  const double *rxx, *rxy, *rxz;
  const int *rix, *riy, *riz;
  int comps[6];              // component ids of this group (field order), ncomp_group entries
  int ncomp_group, comp_base;  // index of the first group component among NComp
  int ncomp, nt;             // NComp, Nt
  int slot_store;            // 0 or 1: ef0 slot written
  double *ef0;               // [slot][2][NComp][nrmax]
  double *efs;               // [slot][nrmax][NComp][Nt]
  int nrmax;
  const int *op_it;          // time-interpolation ops of this event
  const double *op_l1;
  int nops;
};

struct FftPlan {
  int n;            // transform length
  int m;            // power-of-two working length (n, or Bluestein length >= 2n-1)
  int log2m;
  const double *tw;      // m complex twiddles exp(-2 pi i k / m), interleaved re,im
  const double *chirp;   // n complex w_k = exp(-i pi k^2 / n) (Bluestein only)
  const double *bhat;    // m complex FFT of the conjugate chirp, scaled by 1/m (Bluestein only)
};

template <typename T>
struct UpcontArgs {
  const T *bz0;            // plane origin (slot 0), row pitch pz
  T *bxair, *byair;        // plane origins (slot 0), row pitches pxa, pya
  long long pstride;
  int L, M, PX, L1, M1, AnLx, AnRx, Any, AL, AY, Nx, Ny;
  int pz, pxa, pya;        // row pitches of bz0, bxair, byair (PX for the context's own planes)
  int n1, n2, n3;          // spline window lengths
  const double *wl, *wr;   // x-window weights, tap-major: n1 x AnLx, n2 x AnRx
  const double *wf, *wb;   // y-window weights: Any x n3 (front, back)
  // Chain selection: tiled = 1 runs the four-kernel tiled chain (both FFT working lengths
  // <= 2048), 0 the six-kernel line chain.
  int tiled;
  // line chain work planes
  double *tx;              // per slot M x L1: Bz(0) resampled along x (irregular rows)
  double *ry;              // per slot M1 x L1: the y-padding rows of the regular plane
  double2 *cplx;           // per slot M1 x L1 complex work (column spectra / row results)
  long long wstride;       // elements between slots of the line-chain work planes
  // tiled chain work planes, in the field precision T (complex entries: T pairs); slot strides
  // in entries of each plane
  void *A;                 // per slot H x MA complex: x half-spectra of the irregular rows, kx-major
  void *G;                 // per slot H x MG pairs (Gx, Gy) of complex: y-inverse columns, kx-major
  T *tbx, *tby;            // per slot M1 rows of pitch XR: Bx, By on the regular plane
  int MA, MG;
  long long sA, sG;
  // line chain results (FP64), rows of pitch XR
  double *bxr, *byr;
  int XR;
  long long sX;            // slot stride of the results of either chain (entries)
  FftPlan px, py;          // x (length L1) and y (length M1) plans
  const double *gx, *gy;   // filter gains on the half spectrum: M1 x (L1/2+1)
  const double *gxT, *gyT; // the same, kx-major (H x M1), with the non-Hermitian bins zeroed
  int H;                   // L1/2+1
  // bilinear (UpContinuation_gpu.cu:42-110)
  const double *X, *Y, *Xn, *Yn, *dx, *dy;
  const int *xBx2N, *yo2N, *xo2N, *yBy2N;
  const int *yrows;        // the nyr regular rows (ascending) that the bilinear step reads
  int nyr;
};

// A launch whose dispatch packet records e0 at its start and e1 at its end when either is given;
// without events a plain launch (the form that a stream capture records).
template <typename F, typename... A>
inline void launch_k(F kern, dim3 grid, dim3 block, std::uint32_t lds, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                     A... args) {
  if (e0 || e1)
    hipExtLaunchKernelGGL(kern, grid, block, lds, s, e0, e1, 0u, args...);
  else
    hipLaunchKernelGGL(kern, grid, block, lds, s, args...);
}

// ---- launchers (tcsem_kernels.hip, tcsem_upcont.hip) ------------------------------------------
// Optional events e0/e1 are recorded by the dispatch packets themselves (start of the first and
// end of the last kernel of the launcher), so timing adds no packets to the stream.
// geo (device array of nslot MarchArgs records, or null): a model batch of uneven grids, each
// slot with its own fields, tables, sizes and tiling; a then carries the step scalars, the level
// range and the largest N, nper and sper over the records.
template <typename T>
hipError_t launch_march(const MarchArgs<T> &a, int nslot, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr, const MarchArgs<T> *geo = nullptr);
// Rows of the two-step march's workgroups that a precision supports (0: none).
int pair_rows_supported(int prec, int rows);
template <typename T>
hipError_t launch_march2(const PairArgs<T> &p, int nslot, hipStream_t s, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr, const PairArgs<T> *geo = nullptr);
template <typename T>
hipError_t launch_surface(const MarchArgs<T> &a, int nslot, hipStream_t s, hipEvent_t e0 = nullptr,
                          hipEvent_t e1 = nullptr, const MarchArgs<T> *geo = nullptr);
template <typename T>
hipError_t launch_zsum(const MarchArgs<T> &a, int nslot, hipStream_t s, hipEvent_t e0 = nullptr);
// The source pass of a source step after the march (geo: one workgroup per model-batch member, else
// per slot); with a.ibw set it moves the chunk sums of B the march just wrote along with the B cells
// it recomputes.
template <typename T>
hipError_t launch_src(const MarchArgs<T> &a, int nslot, hipStream_t s, const MarchArgs<T> *geo = nullptr);
template <typename T>
hipError_t launch_bz_volume(const MarchArgs<T> &a, T *bz, int slot, hipStream_t s);
template <typename T>
hipError_t launch_gather(const GatherArgs &g, int is_b, int nslot, const MarchArgs<T> &a,
                         hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// geo (device array of nslot UpcontArgs records, or null): the air planes of a model batch; u then
// carries the largest L, M, L1, M1, H and FFT working lengths over the records (tiled chain only).
// alone: the chain runs with nothing beside it on the GPU (step A's chain in a pair of steps), so
// the tiled chain takes 256 threads per line instead of the shape sized to share the CUs with a
// concurrent march.  The results do not depend on it.
template <typename T>
hipError_t launch_upcont(const UpcontArgs<T> &u, int nslot, hipStream_t s, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr, const UpcontArgs<T> *geo = nullptr, bool alone = false);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);
// Work planes of the allocation-free upward continuation: a file-scope device array of
// kUpcontScratchBytes per device (tcsem_upcont.hip).  upcont_scratch_init looks its address up on
// device dev (current) once; upcont_scratch returns that address (null before the look-up).
constexpr size_t kUpcontScratchBytes = (size_t)40 << 20;
hipError_t upcont_scratch_init(int dev);
void *upcont_scratch(int dev);
// Step scalars of a graph replay: dst[0 .. 4 nsteps) = the values (lam, dt, dtb, amp per step).
constexpr int kGraphMaxSteps = 32;
// Contexts of fewer cells start with graph replays of kGraphAutoSteps steps (their steps are bound
// by the host's launches, not by the device).
constexpr long long kGraphAutoCells = 1LL << 21;
constexpr int kGraphAutoSteps = 16;
struct GraphScalars {
  double v[4 * kGraphMaxSteps];
};
hipError_t launch_put_scalars(double *dst, const GraphScalars &s, int nsteps, hipStream_t st);

}  // namespace tcsem
