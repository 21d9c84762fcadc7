/*
 * tcsem.h -- C-ABI of the MI355X-native TCSEM_FDTD time-stepping library (libtcsem.so).
 *
 * The library advances the DuFort-Frankel scheme of xltu/TCSEM_FDTD on one AMD
 * Instinct MI355X (gfx950): E update, source injection, B update with the
 * div(B)=0 Bz integral, FFT upward continuation of Bz at the sea surface, and
 * receiver sampling.  It reads and writes the reference's own driver files.
 *
 * Every entry point returns 0 on success or a negative TCSEM_E* code; none of
 * them exits the process.  tcsem_last_error() returns a message for the last
 * failure on the calling thread.  Arrays are plain host pointers owned by the
 * caller; a context owns all device memory.  One context is driven by one host
 * thread at a time.
 *
 * Reference interfaces each group replaces (file:line in xltu/TCSEM_FDTD):
 *   file surface     FDTD_setup.in  ModelingSetup.cpp:51-59
 *                    Cond.bin       ModelingSetup.cpp:201-209
 *                    Receiver.in    LoadRxTxPos.cpp:51-84
 *                    Transmitter.in LoadRxTxPos.cpp:134-200
 *                    Tx<id>.bout    main.cu:280-291
 *   host setup       ModelingSetup.cpp:100-192, ModelGridMapping.cpp:13-395,
 *                    SourceWave.cpp:8-58, LoadRxTxPos.cpp:93-460, main.cu:149-165
 *   device setup     prepare_free_mem.cu:45-239, initialize_UpCont_gpu_resources.cu:146-341,
 *                    initialize_fft.cu:43-156
 *   step API         UpContinuation_gpu  FTDT.h:317-318 (UpContinuation_gpu.cu:166-274)
 *                    ElectricalField     FTDT.h:296     (ElectricalFiled.cu:478-655)
 *                    Bderivate           FTDT.h:291     (Bderivate.cu:269-316)
 *                    EFInterp2Rx         FTDT.h:342-344 (EFInterp2Rx.cu:472-585)
 *   driver loop      main.cu:195-298
 */
#ifndef TCSEM_H
#define TCSEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------------------------ */
#define TCSEM_OK 0
#define TCSEM_EINVAL (-1)     /* bad argument or inconsistent sizes                           */
#define TCSEM_EDOMAIN (-2)    /* receiver / transmitter outside the model                     */
#define TCSEM_EIO (-3)        /* file open / parse / write failure                            */
#define TCSEM_ENOMEM (-4)     /* host or device allocation failed                             */
#define TCSEM_EDEVICE (-5)    /* HIP runtime error or no usable gfx950 device                 */
#define TCSEM_ESIZE (-6)      /* size beyond a supported limit                                */
#define TCSEM_ESTATE (-7)     /* call out of order (e.g. run before transmitters were set)    */

/* ---- precision of the field state --------------------------------------------------------- */
#define TCSEM_FP64 64         /* fields, coefficients and sums in FP64 (reference precision)  */
#define TCSEM_FP32 32         /* FP32 fields and air-plane FFT (tables rounded from FP64); FP64
                                 coefficients, schedule, receiver splines; compensated FP32 Bz
                                 integral                                                      */

/* ---- field component ids (tcsem_get_field / tcsem_set_field) ------------------------------ */
#define TCSEM_EX 0            /* (N+1, M+1, L)   x fastest */
#define TCSEM_EY 1            /* (N+1, M,   L+1) */
#define TCSEM_EZ 2            /* (N,   M+1, L+1) */
#define TCSEM_BX 3            /* (N,   M,   L+1) */
#define TCSEM_BY 4            /* (N,   M+1, L)   */
#define TCSEM_BZ 5            /* (N+1, M,   L)   derived: bottom-up div(B)=0 integral of Bx, By */
#define TCSEM_BXAIR 6         /* (M,   L+1)      Bx at z=-dz0/2 from the last upward continuation */
#define TCSEM_BYAIR 7         /* (M+1, L)        By at z=-dz0/2 */

/* FDTD_setup.in, in file order (ModelingSetup.cpp:51-59). */
typedef struct tcsem_setup {
  double sfa1, sfa2;                 /* line 1: time-step speed-up factor, lambda stabiliser   */
  int n_src;                         /* line 2: Gaussian discretisation count (N_Src_Steps = n_src/2) */
  double sigma_t;                    /* line 2: Gaussian width [s]                             */
  double x0, y0, z0;                 /* line 3: grid origin                                    */
  double dxmin, dymin, dzmin;        /* line 4: core cell sizes [m]                            */
  double arg1, arg2;                 /* line 5: geometric growth factors                       */
  int ALx2, ALx1, Nx, ARx1, ARx2;    /* line 6: x zone cell counts, outer-left first           */
  int Ay2, Ay1, Ny;                  /* line 7: y zones (symmetric)                            */
  int Az2, Az1, Nz;                  /* line 8: z zones (file order Az2 Az1 Nz)                */
} tcsem_setup;

/* Receiver.in lines 1-3 (LoadRxTxPos.cpp:51-55). */
typedef struct tcsem_rxtime {
  double T0, Tn;                     /* recording window [s]                                   */
  int Nt;                            /* samples per component                                  */
  int is_log;                        /* 1: log-spaced sample times, 0: linear                  */
  int is_comp[6];                    /* record Ex Ey Ez Bx By Bz                               */
} tcsem_rxtime;

/* One Transmitter.in row (LoadRxTxPos.cpp:151-178).  ort 1-3: x/y/z bipole; ort 4: one bipole
 * with azimuth/dip; ort 5: two bipoles.  Angles are in degrees as in the file.  rx_ids holds nr
 * 1-based receiver ids that record this transmitter. */
typedef struct tcsem_tx {
  int id, ort;
  double bipolen, x, y, z, azi_deg, dip_deg;
  double bipolen1, x1, y1, z1, azi1_deg, dip1_deg;
  int nr;
  const int *rx_ids;
} tcsem_tx;

/* Sizes derived from a setup (ModelingSetup.cpp:79-84,190-192; main.cu:102-103). */
typedef struct tcsem_grid_info {
  int L, M, N;                       /* cells in x, y, z                                       */
  int AnLx_New, AnRx_New, Any_New;   /* regular-grid padding widths of the air plane           */
  int L1, M1;                        /* regular air-plane size (x, y)                          */
} tcsem_grid_info;

typedef struct tcsem_ctx tcsem_ctx;

/* ---- library ------------------------------------------------------------------------------- */
const char *tcsem_version(void);
const char *tcsem_last_error(void);

/* ---- file surface (host only, no device use) ---------------------------------------------- */
/* ModelingSetup.cpp:43-61 */
int tcsem_read_setup(const char *path, tcsem_setup *out);
/* ModelingSetup.cpp:201-209: L*M*N FP64, x fastest. */
int tcsem_read_cond(const char *path, size_t ncell, double *out);
/* LoadRxTxPos.cpp:43-86.  Pass xyz=NULL to query *nrx; otherwise xyz holds 3*cap doubles
 * (x,y,z per receiver, indexed by id-1). */
int tcsem_read_receivers(const char *path, tcsem_rxtime *rt, int *nrx, double *xyz, int cap);
/* LoadRxTxPos.cpp:126-200.  Pass out=NULL to query *ntx and *nids (total receiver ids); then
 * out[cap] rows are filled and their rx_ids point into ids[ids_cap]. */
int tcsem_read_transmitters(const char *path, tcsem_tx *out, int cap, int *ntx, int *ids,
                            int ids_cap, int *nids);
/* main.cu:280-291: int32 {nr, ncomp, nt} then FP64 [nr][ncomp][nt]. */
int tcsem_write_bout(const char *path, int nr, int ncomp, int nt, const double *data);
int tcsem_read_bout(const char *path, int *nr, int *ncomp, int *nt, double *data, size_t cap);

/* ---- host setup restatements (no device use) ---------------------------------------------- */
int tcsem_grid_info_of(const tcsem_setup *s, tcsem_grid_info *out);
/* ModelingSetup.cpp:100-187, ModelGridMapping.cpp:58-75.  Any pointer may be NULL. */
int tcsem_grid_arrays(const tcsem_setup *s, double *dx, double *dy, double *dz, double *X,
                      double *Y, double *Z, double *Xnew, double *Ynew);
/* ModelGridMapping.cpp:77-366.  Any pointer may be NULL. */
int tcsem_grid_maps(const tcsem_setup *s, int *xN2o, int *yN2o, int *xo2N, int *yo2N,
                    int *xBx2N, int *yBy2N);
/* SourceWave.cpp:8-58 (t, wave: 2*ns+1 values; dt: 2*ns+2 values). */
int tcsem_waveform(int ns, double sigma_t, double *t, double *dt, double *wave);

/* ---- device context ------------------------------------------------------------------------ */
/* ModelingSetup + ModelGridMapping + prepare_device_arrays + initialize_UpCont_gpu_resources.
 * cond: L*M*N FP64 conductivity, x fastest.  precision: TCSEM_FP64 or TCSEM_FP32. */
int tcsem_create(const tcsem_setup *s, const double *cond, int precision, tcsem_ctx **out);
void tcsem_destroy(tcsem_ctx *ctx);
int tcsem_info(const tcsem_ctx *ctx, tcsem_grid_info *out);

/* LoadRxTxPos.cpp:51-121,431-460.  rx_xyz: 3*nrx (x,y,z per receiver id-1). */
int tcsem_set_receivers(tcsem_ctx *ctx, const tcsem_rxtime *rt, int nrx, const double *rx_xyz);
/* LoadRxTxPos.cpp:143-428 (index search) and ElectricalFiled.cu:532-641 (bipole currents). */
int tcsem_set_transmitters(tcsem_ctx *ctx, const tcsem_tx *tx, int ntx);

/* Number of field sets (transmitters advanced together) the context holds; grows on demand. */
int tcsem_reserve_batch(tcsem_ctx *ctx, int nbatch);
/* set_zeros_EM_arrays (prepare_free_mem.cu:164-186) for field sets [first, first+n). */
int tcsem_reset_fields(tcsem_ctx *ctx, int first, int n);

/* Autotuner of the volume march (no counterpart in the reference, whose launch shapes are
 * fixed): times nsteps nbatch-transmitter steps as they run in production (the march beside the
 * upward continuation and surface level on the second stream; no source, no gathers) for every
 * tiling the library runs -- rows per workgroup 4/8/10/16 (FP64: 4/8/10/12), prefetch depth 1-4,
 * z-chunks of the lower launch 1/2/3/4/6/8 as the depth allows, for chunked tilings the source of
 * the chunks' Bz partial sums (a pass over Bx, By, or the chunk sums the previous step's march
 * accumulated), and the 8-bit conductivity index on and off -- then, with that tiling for the top
 * levels, pairs of steps in the two-step pass (rows 8/10/12/14/16, FP64 8/10/11/12, prefetch 1-2; tcsem_set_pairs;
 * every shape timed in two interleaved rounds, the faster kept), kept when 2% faster per step -- and keeps the choice for later batches of that size
 * (TCSEM_MARCH_ROWS / _PF / _ZCHUNKS / _ZSUM / _PAIRS and tcsem_set_pairs still override it).  It runs on the fields of slots [0, nbatch): a begun run
 * is discarded and the fields are left zeroed.  Outputs (nullable): the chosen rows, prefetch
 * depth, z-chunks, partial-sum source (1 chunk sums, 0 pass) and the time per step [ms]. */
int tcsem_autotune(tcsem_ctx *ctx, int nbatch, int nsteps, int *rows, int *pf, int *zchunks, int *zsum,
                   double *ms);

/* One DuFort-Frankel step for field sets [0, nbatch), each driven by transmitter
 * tx_first+b with the schedule of the step counter.  Equivalent to main.cu:223-260:
 * UpContinuation_gpu, ElectricalField (with source), Bderivate, then time advance. */
int tcsem_begin(tcsem_ctx *ctx, int tx_first, int nbatch);
int tcsem_step(tcsem_ctx *ctx, int nsteps);
/* Steps still needed until every receiver sample is taken (main.cu:219). */
int tcsem_steps_remaining(const tcsem_ctx *ctx);

/* Full per-transmitter run (main.cu:195-298) for transmitters [tx_first, tx_first+ntx), all
 * advanced in one batch.  traces: ntx blocks of [nr][NComp][Nt] FP64 (nr of each Tx).
 * max_steps > 0 caps the loop.  *steps_done (nullable) receives the steps taken. */
int tcsem_run(tcsem_ctx *ctx, int tx_first, int ntx, int max_steps, double *traces,
              int *steps_done);
/* Traces of batch slot b after tcsem_begin/tcsem_step (nr x NComp x Nt FP64, nr of transmitter
 * tx_first+slot of the begun run).  The copy into efs of main.cu:278 (d_efs of xyzTx[itx].nr
 * receivers, main.cu:207). */
int tcsem_get_traces(tcsem_ctx *ctx, int slot, double *traces);
/* Shape of tcsem_get_traces' output for batch slot `slot` of the begun run: receivers of that
 * slot's transmitter, recorded components, samples (the reference sizes its efs from the same
 * numbers, main.cu:207-210).  Callers size their buffer from it. */
int tcsem_trace_shape(const tcsem_ctx *ctx, int slot, int *nr, int *ncomp, int *nt);

/* Field access for batch slot `slot`; FP32 contexts convert.  Sizes as the component ids. */
int tcsem_get_field(tcsem_ctx *ctx, int slot, int comp, double *out);
int tcsem_set_field(tcsem_ctx *ctx, int slot, int comp, const double *in);

/* UpContinuation_gpu on a caller plane: bz0 (M*L) -> bxair (M*(L+1)), byair ((M+1)*L). */
int tcsem_upcont_plane(tcsem_ctx *ctx, const double *bz0, double *bxair, double *byair);

/* Allocation-free form of tcsem_upcont_plane (no counterpart in the reference): for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same kernels and same results as tcsem_upcont_plane; the planes are the
 * caller's device buffers in the context's field precision (double for TCSEM_FP64, float for
 * TCSEM_FP32), dense rows: d_bz0 M*L, d_bxair M*(L+1), d_byair (M+1)*L.  The work planes live in
 * a fixed device array of the library (40 MiB per device, held until the process ends); air
 * planes whose work exceeds it, or whose FFT lengths exceed 2048, return TCSEM_ESIZE.  The call
 * runs on the context's stream, allocates nothing, and returns once the planes are written.  All
 * calls on one device share that array: calls that overlap in time on one device (from several
 * host threads) race. */
int tcsem_upcont_plane_noalloc(tcsem_ctx *ctx, const void *d_bz0, void *d_bxair, void *d_byair);

/* Average device time [ms] of the kernels of the steps taken while timing was enabled, from HIP
 * events that the kernels' dispatch packets record on the context's streams as the steps run
 * (every 8th step, or the pair of steps holding it; no extra synchronisation, reading waits for
 * them): t[0] volume march per step (lower and top launch; the lower from the start of its
 * k_zsum pass when the step runs one; for a pair of steps half of the two-step pass and both top
 * launches), t[1] surface level, t[2] upward continuation chain, t[3] receiver gather per launch;
 * n[i] counts. */
int tcsem_timing(const tcsem_ctx *ctx, double t[4], long n[4]);
int tcsem_enable_timing(tcsem_ctx *ctx, int on);
int tcsem_sync(tcsem_ctx *ctx);

/* Graph replays (no counterpart in the reference, which launches every kernel of every step):
 * with nsteps > 0 (even, at most 32), tcsem_step and tcsem_run run each stretch of nsteps
 * steps after the source wavelet that holds no receiver sample as one replay of a HIP graph.
 * The graph is captured on first use for its starting state; the step scalars are written to
 * device memory before each replay.  Results equal those of the plain steps (tested).  It pays
 * where the host's launches bound the step (small grids): contexts of fewer than 2^21 cells start
 * with 16, larger ones with 0 (off); 0 turns it off.
 * tcsem_graph_replays counts the replays of the context so far. */
int tcsem_set_graph(tcsem_ctx *ctx, int nsteps);
long tcsem_graph_replays(const tcsem_ctx *ctx);

/* The conductivity as the volume march reads it (no counterpart in the reference): a model of at
 * most 255 distinct nonzero conductivities (as the field precision holds them) is read as an
 * 8-bit index into a table of those values, one byte per cell instead of a word, with the same
 * results; TCSEM_SIGMA8=0 in the environment at tcsem_create turns it off, and tcsem_autotune
 * times the march both ways and keeps the faster.  Returns the table's entries in use (0
 * included), 0 when the march reads the full conductivity, -1 for no context. */
int tcsem_sigma_table(const tcsem_ctx *ctx);

/* Two steps per lower pass (no counterpart in the reference, whose steps each read and write
 * the whole state): with rows 8, 10, 12, 14 or 16 (FP64: 8, 10, 11 or 12; workgroup rows of the pass) and pf 1 or 2 (prefetch
 * depth), tcsem_step and tcsem_run take each two consecutive steps after the source wavelet
 * whose first holds no receiver sample as one pass over levels N-1 .. 3 of both steps (the
 * intermediate state stays on chip there), then the top levels of each step as usual.  The
 * results equal, bit for bit, those of one step at a time with an unchunked march (tested).
 * rows 0 turns it off.  Without a call the library decides at tcsem_begin: TCSEM_MARCH_PAIRS ("0"
 * or "rows,pf"), else the autotuner's choice for the batch size, else 12-row passes where they
 * give at least two workgroups per CU (14-row FP32 or 11-row FP64 passes where those leave fewer
 * idle CUs in the last round of workgroups); a call holds for the context's later runs.  Graph replays,
 * where on, take precedence.  Needs N >= 6.
 * tcsem_pair_info reports the setting and the pairs taken so far; tcsem_pair_timing the average
 * device time [ms] of the two-step pass alone (timing enabled; its top launches count in
 * tcsem_timing's t[0], per step). */
int tcsem_set_pairs(tcsem_ctx *ctx, int rows, int pf);
int tcsem_pair_info(const tcsem_ctx *ctx, int *rows, int *pf, long *pairs);
int tcsem_pair_timing(const tcsem_ctx *ctx, double *ms, long *n);

/* ---- model batch (no counterpart in the reference, which runs one model per process) --------
 * Conductivity models of uneven grids, one transmitter each, advanced together: every step's
 * volume march, air-plane chain and surface level run as one launch over all members.  Each
 * member is a context of its own (tcsem_create, tcsem_set_receivers, tcsem_set_transmitters),
 * begun with one transmitter (tcsem_begin(ctx, tx, 1)); all share the precision, have N >= 6
 * levels and air planes whose FFT lengths are at most 2048.  Members stay owned by the caller
 * and must outlive the batch; their traces are read with tcsem_get_traces(ctx, 0, ...) as
 * usual.  A member whose receivers have all been sampled is left out of later calls.
 *   tcsem_batch_autotune: times nsteps batch steps for every rows x prefetch tiling (before the
 *     first step; the members' fields are zeroed again) and keeps the fastest, then pairs of
 *     steps in the two-step pass (rows 8/10/12/14/16, FP64 8/10/11/12; two interleaved rounds), kept
 *     when 2% faster per step.
 *   tcsem_batch_step: nsteps steps of every member still sampling (per member as tcsem_step).
 *   tcsem_batch_run: steps until every member has taken all samples (max_steps > 0 caps the
 *     member step count; members finish at the end of a 256-step chunk). */
typedef struct tcsem_batch tcsem_batch;
int tcsem_batch_create(tcsem_ctx *const *models, int nmodel, tcsem_batch **out);
int tcsem_batch_autotune(tcsem_batch *batch, int nsteps, int *rows, int *pf, double *ms);
int tcsem_batch_step(tcsem_batch *batch, int nsteps);
int tcsem_batch_run(tcsem_batch *batch, int max_steps, int *steps_done);
int tcsem_batch_sync(tcsem_batch *batch);
/* As tcsem_timing for the batch's launches: t[0] volume march per step (lower and top launch
 * over all members; for a pair of steps half of the two-step pass and both top launches), t[1]
 * surface level, t[2] air-plane chain; n[i] counts. */
int tcsem_batch_enable_timing(tcsem_batch *batch, int on);
int tcsem_batch_timing(tcsem_batch *batch, double t[3], long n[3]);
/* Two steps per pass for the batch (no counterpart in the reference), as tcsem_set_pairs for a
 * context: rows 8, 10, 12, 14 or 16 (FP64: 8, 10, 11 or 12) at prefetch depth 1, or 0 (one step at a time), over
 * the library's choice (TCSEM_MARCH_PAIRS, the batch autotuner's pick, else 12-row passes when
 * the members' pass tiles give at least two workgroups per CU).  Pairs run when every member
 * still sampling is past its source wavelet and takes no sample after the pair's first step.
 * The results equal one step at a time, bit for bit (tested).  tcsem_batch_pair_info: the rows
 * of the last step call (0: one step at a time), the autotuner's pick (-1: not tuned) and the
 * pairs taken; tcsem_batch_pair_timing: the pass's average device time while timing is enabled. */
int tcsem_batch_set_pairs(tcsem_batch *batch, int rows);
int tcsem_batch_pair_info(const tcsem_batch *batch, int *rows, int *tuned, long *pairs);
int tcsem_batch_pair_timing(tcsem_batch *batch, double *ms, long *n);
void tcsem_batch_destroy(tcsem_batch *batch);

#ifdef __cplusplus
}
#endif
#endif /* TCSEM_H */
