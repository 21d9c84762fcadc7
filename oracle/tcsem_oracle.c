/*
 * tcsem_oracle.c -- FP64 CPU restatement of the xltu/TCSEM_FDTD time-stepping hot path.
 *
 * TEST INFRASTRUCTURE.  This is the checker, not the product: only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py load it (through oracle/oracle.py).
 * The product library (tcsem_fdtd_amd/csrc) never includes, links or calls anything here.
 *
 * It restates, loop for loop, the CUDA formulation of the reference (the only formulation
 * that ran): grid and mapping setup, the Gaussian source schedule, DuFort-Frankel E update
 * with bipole injection, curl-E B update and the bottom-up div(B)=0 Bz integral, the GPU-form
 * natural-spline IR->R resampling, the FFT wavenumber filter, bilinear R->IR, the 64-point
 * receiver gathers and the host time-interpolation state machine.  Each function cites the
 * reference file:line it follows.  It implements the intended, race-free semantics (every
 * level reads same-step values).  The FFT is a plain radix-2 / Bluestein FFT written here
 * (FFTW/cuFFT are absent); tests pin it against numpy.fft.  It also restates the reference's CPU
 * twins of the air boundary (SplineGridConvIR2R_cpu.cpp, the filter of UpContinuation.cpp,
 * BilinearGridConvR2IR.cpp), selectable for the air boundary of a step, to compare the two
 * formulations.
 *
 * Pins (see tests/test_oracle_*.py): testwaveform.dat (golden), closed-form grid invariants,
 * the reference's DEBUG==1 mapping bracket checks, scipy natural CubicSpline, numpy.fft,
 * exactness of bilinear / 4-point splines on linear data, div(B)=0, linearity, mirror
 * symmetry, the homogeneous DF decay factor, and whole-space/half-space closed forms; the CPU
 * twins against scipy natural splines on their own node spacing and against the GPU formulation
 * where the two coincide (filter with dx = dy, bilinear).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/tcsem.h"

#define ORC_PI 3.141592653589793238 /* FTDT.h:39 */
#define ORC_MIU (4 * ORC_PI * 1.e-7) /* main.cu:55, ElectricalFiled.cu:33 */
#define ORC_NPAD 4                  /* FTDT.h:40 */

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* ============================================================================================
 * model
 * ========================================================================================== */
typedef struct orc_model {
  tcsem_setup s;
  int L, M, N, AnLx, AnRx, Any, L1, M1, AL, AR, AY;
  double mindx;
  double *dx, *dy, *dz, *X, *Y, *Z, *Xn, *Yn;
  int *xN2o, *yN2o, *xo2N, *yo2N, *xBx2N, *yBy2N;
  double *cond;
  double *Ex, *Ey, *Ez, *Bx, *By, *Bz, *BxAir, *ByAir;
  double *T, *BzAir, *Fr, *Fi, *Gr, *Gi; /* up-continuation work */
  int cpu_twins; /* 1: the air boundary of a step uses the CPU-twin chain */
} orc_model;

/* ModelingSetup.cpp:79-84,100-192 */
static int orc_build_grid(orc_model *m) {
  tcsem_setup *s = &m->s;
  int i, j, k, nc;
  m->AL = s->ALx1 + s->ALx2;
  m->AR = s->ARx1 + s->ARx2;
  m->AY = s->Ay1 + s->Ay2;
  m->L = s->ALx2 + s->ALx1 + s->ARx2 + s->ARx1 + s->Nx;
  m->M = 2 * (s->Ay2 + s->Ay1) + s->Ny;
  m->N = s->Nz + s->Az2 + s->Az1;
  int L = m->L, M = m->M, N = m->N;
  /* zero padding indexes X_Bzold[-1] in the reference (ModelingSetup.cpp:190-192) */
  if (m->AL < 1 || m->AR < 1 || m->AY < 1 || s->Nx < ORC_NPAD || s->Ny < ORC_NPAD || N < 3)
    return TCSEM_EINVAL;
  m->dx = calloc(L, sizeof(double));
  m->dy = calloc(M, sizeof(double));
  m->dz = calloc(N, sizeof(double));
  m->X = calloc(L, sizeof(double));
  m->Y = calloc(M, sizeof(double));
  m->Z = calloc(N, sizeof(double));
  for (i = m->AL; i < s->Nx + m->AL; i++) m->dx[i] = s->dxmin;
  for (i = m->AL - 1; i >= s->ALx2; i--) m->dx[i] = s->dxmin * pow(s->arg1, (double)(m->AL - i));
  for (i = s->ALx2 - 1; i >= 0; i--)
    m->dx[i] = s->dxmin * pow(s->arg1, (double)s->ALx1) * pow(s->arg2, (double)(s->ALx2 - i));
  for (i = s->Nx + m->AL; i < s->Nx + m->AL + s->ARx1; i++)
    m->dx[i] = s->dxmin * pow(s->arg1, (double)(i - m->AL - s->Nx + 1));
  for (i = s->Nx + m->AL + s->ARx1; i < L; i++)
    m->dx[i] = s->dxmin * pow(s->arg1, (double)s->ARx1) *
               pow(s->arg2, (double)(i - s->Nx - m->AL - s->ARx1 + 1));
  for (j = m->AY; j < s->Ny + m->AY; j++) m->dy[j] = s->dymin;
  for (j = m->AY - 1; j >= s->Ay2; j--) m->dy[j] = m->dy[j + 1] * s->arg1;
  for (j = s->Ay2 - 1; j >= 0; j--) m->dy[j] = m->dy[j + 1] * s->arg2;
  for (j = s->Ny + m->AY; j < s->Ny + s->Ay2 + 2 * s->Ay1; j++) m->dy[j] = m->dy[j - 1] * s->arg1;
  for (j = s->Ny + s->Ay2 + 2 * s->Ay1; j < M; j++) m->dy[j] = m->dy[j - 1] * s->arg2;
  for (k = 0; k < s->Nz; k++) m->dz[k] = s->dzmin;
  for (k = s->Nz; k < s->Nz + s->Az1; k++) m->dz[k] = m->dz[k - 1] * s->arg1;
  for (k = s->Nz + s->Az1; k < N; k++) m->dz[k] = m->dz[k - 1] * s->arg2;
  if (s->Nx % 2 == 0) {
    nc = s->Nx / 2 + m->AL;
    m->X[nc] = 0.5 * m->dx[nc] + s->x0;
  } else {
    nc = (s->Nx - 1) / 2 + m->AL;
    m->X[nc] = s->x0;
  }
  for (i = nc + 1; i < L; i++) m->X[i] = m->X[i - 1] + 0.5 * m->dx[i - 1] + 0.5 * m->dx[i];
  for (i = nc - 1; i >= 0; i--) m->X[i] = m->X[i + 1] - 0.5 * m->dx[i + 1] - 0.5 * m->dx[i];
  if (s->Ny % 2 == 0) {
    nc = s->Ny / 2 + m->AY;
    m->Y[nc] = 0.5 * m->dy[nc] + s->y0;
  } else {
    nc = (s->Ny - 1) / 2 + m->AY;
    m->Y[nc] = s->y0;
  }
  for (j = nc + 1; j < M; j++) m->Y[j] = m->Y[j - 1] + 0.5 * m->dy[j - 1] + 0.5 * m->dy[j];
  for (j = nc - 1; j >= 0; j--) m->Y[j] = m->Y[j + 1] - 0.5 * m->dy[j + 1] - 0.5 * m->dy[j];
  m->Z[0] = 0.5 * m->dz[0] + s->z0;
  for (k = 1; k < N; k++) m->Z[k] = m->Z[k - 1] + 0.5 * m->dz[k - 1] + 0.5 * m->dz[k];
  m->AnLx = (int)ceil((m->X[m->AL - 1] - m->X[0] + m->dx[m->AL - 1] / 2 + m->dx[0] / 2) / s->dxmin);
  m->AnRx = (int)ceil((m->X[L - 1] - m->X[L - m->AR] + m->dx[L - 1] / 2 + m->dx[L - m->AR] / 2) /
                      s->dxmin);
  m->Any = (int)ceil((m->Y[m->AY - 1] - m->Y[0] + m->dy[m->AY - 1] / 2 + m->dy[0] / 2) / s->dymin);
  m->L1 = m->AnLx + m->AnRx + s->Nx; /* main.cu:102-103 */
  m->M1 = 2 * m->Any + s->Ny;
  m->mindx = fmin(s->dzmin, fmin(s->dxmin, s->dymin)); /* main.cu:107 */
  return TCSEM_OK;
}

/* ModelGridMapping.cpp:58-366 */
static void orc_build_mapping(orc_model *m) {
  int i, j;
  int L0 = m->L, M0 = m->M, L1 = m->L1, M1 = m->M1, Nx = m->s.Nx, Ny = m->s.Ny;
  double dxmin = m->s.dxmin, dymin = m->s.dymin;
  double *X = m->X, *Y = m->Y, *dx = m->dx, *dy = m->dy;
  m->Xn = calloc(L1, sizeof(double));
  m->Yn = calloc(M1, sizeof(double));
  m->xN2o = calloc(L1, sizeof(int));
  m->yN2o = calloc(M1, sizeof(int));
  m->xo2N = calloc(L0, sizeof(int));
  m->yo2N = calloc(M0, sizeof(int));
  m->xBx2N = calloc(L0 + 1, sizeof(int));
  m->yBy2N = calloc(M0 + 1, sizeof(int));
  double *Xn = m->Xn, *Yn = m->Yn;
  for (i = m->AnLx; i < m->AnLx + Nx; i++) Xn[i] = X[i - m->AnLx + m->AL];
  for (i = m->AnLx - 1; i >= 0; i--) Xn[i] = Xn[i + 1] - dxmin;
  for (i = m->AnLx + Nx; i < L1; i++) Xn[i] = Xn[i - 1] + dxmin;
  for (j = m->Any; j < m->Any + Ny; j++) Yn[j] = Y[j - m->Any + m->AY];
  for (j = m->Any - 1; j >= 0; j--) Yn[j] = Yn[j + 1] - dymin;
  for (j = m->Any + Ny; j < M1; j++) Yn[j] = Yn[j - 1] + dymin;

  int *a = m->xN2o;
  for (i = m->AnLx; i < m->AnLx + Nx; i++) a[i] = i - m->AnLx + m->AL;
  for (i = m->AnLx - 1; i >= 0; i--) {
    j = a[i + 1];
    while (j > 0 && X[j - 1] >= Xn[i]) j--;
    a[i] = j;
  }
  for (i = m->AnLx + Nx; i < L1; i++) {
    j = a[i - 1] - 1;
    while (j < L0 - 1 && X[j + 1] < Xn[i]) j++;
    a[i] = j + 1;
  }
  a = m->yN2o;
  for (j = m->Any; j < m->Any + Ny; j++) a[j] = j - m->Any + m->AY;
  for (j = m->Any - 1; j >= 0; j--) {
    i = a[j + 1];
    while (i > 0 && Y[i - 1] >= Yn[j]) i--;
    a[j] = i;
  }
  for (j = m->Any + Ny; j < M1; j++) {
    i = a[j - 1] - 1;
    while (i < M0 - 1 && Y[i + 1] < Yn[j]) i++;
    a[j] = i + 1;
  }
  a = m->xo2N;
  for (i = m->AL; i < m->AL + Nx; i++) a[i] = i - m->AL + m->AnLx;
  for (i = m->AL - 1; i >= 0; i--) {
    j = a[i + 1];
    while (j > 0 && Xn[j - 1] >= X[i]) j--;
    a[i] = j;
  }
  for (i = m->AL + Nx; i < L0; i++) {
    j = a[i - 1];
    while (j < L1 - 1 && Xn[j + 1] < X[i]) j++;
    a[i] = j + 1;
  }
  a = m->yo2N;
  for (j = m->AY; j < m->AY + Ny; j++) a[j] = j - m->AY + m->Any;
  for (j = m->AY - 1; j >= 0; j--) {
    i = a[j + 1];
    while (i > 0 && Yn[i - 1] >= Y[j]) i--;
    a[j] = i;
  }
  for (j = m->AY + Ny; j < M0; j++) {
    i = a[j - 1];
    while (i < M1 - 1 && Yn[i + 1] < Y[j]) i++;
    a[j] = i + 1;
  }
  a = m->xBx2N;
  for (i = m->AL; i < m->AL + Nx; i++) a[i] = i - m->AL + m->AnLx;
  for (i = m->AL - 1; i >= 0; i--) {
    j = a[i + 1];
    while (j > 0 && Xn[j - 1] >= X[i] - 0.5 * dx[i]) j--;
    a[i] = j;
  }
  for (i = m->AL + Nx; i < L0; i++) {
    j = a[i - 1];
    while (j < L1 - 1 && Xn[j + 1] < X[i] - 0.5 * dx[i]) j++;
    a[i] = j + 1;
  }
  i = L0;
  j = a[i - 1];
  while (j < L1 - 1 && Xn[j + 1] < X[i - 1] + 0.5 * dx[i - 1]) j++;
  a[i] = j + 1;
  a = m->yBy2N;
  for (j = m->AY; j < m->AY + Ny; j++) a[j] = j - m->AY + m->Any;
  for (j = m->AY - 1; j >= 0; j--) {
    i = a[j + 1];
    while (i > 0 && Yn[i - 1] >= Y[j] - 0.5 * dy[j]) i--;
    a[j] = i;
  }
  for (j = m->AY + Ny; j < M0; j++) {
    i = a[j - 1];
    while (i < M1 - 1 && Yn[i + 1] < Y[j] - 0.5 * dy[j]) i++;
    a[j] = i + 1;
  }
  j = M0;
  i = a[j - 1];
  while (i < M1 - 1 && Yn[i + 1] < Y[j - 1] + 0.5 * dy[j - 1]) i++;
  a[j] = i + 1;
}

#define SZ_EX(m) ((size_t)((m)->N + 1) * ((m)->M + 1) * (m)->L)
#define SZ_EY(m) ((size_t)((m)->N + 1) * (m)->M * ((m)->L + 1))
#define SZ_EZ(m) ((size_t)(m)->N * ((m)->M + 1) * ((m)->L + 1))
#define SZ_BX(m) ((size_t)(m)->N * (m)->M * ((m)->L + 1))
#define SZ_BY(m) ((size_t)(m)->N * ((m)->M + 1) * (m)->L)
#define SZ_BZ(m) ((size_t)((m)->N + 1) * (m)->M * (m)->L)

void orc_destroy(orc_model *m);

/* dx/dy/dz may override the setup's spacing (NULL = reference formulas); X/Y/Z are then rebuilt
 * with the same recurrences. */
orc_model *orc_create(const tcsem_setup *s, const double *cond, const double *dxo,
                      const double *dyo, const double *dzo) {
  orc_model *m = calloc(1, sizeof(orc_model));
  m->s = *s;
  if (orc_build_grid(m) != TCSEM_OK) {
    orc_destroy(m);
    return NULL;
  }
  if (dxo || dyo || dzo) {
    int i, j, k, nc;
    if (dxo) memcpy(m->dx, dxo, m->L * sizeof(double));
    if (dyo) memcpy(m->dy, dyo, m->M * sizeof(double));
    if (dzo) memcpy(m->dz, dzo, m->N * sizeof(double));
    nc = (s->Nx % 2 == 0) ? s->Nx / 2 + m->AL : (s->Nx - 1) / 2 + m->AL;
    m->X[nc] = (s->Nx % 2 == 0) ? 0.5 * m->dx[nc] + s->x0 : s->x0;
    for (i = nc + 1; i < m->L; i++) m->X[i] = m->X[i - 1] + 0.5 * m->dx[i - 1] + 0.5 * m->dx[i];
    for (i = nc - 1; i >= 0; i--) m->X[i] = m->X[i + 1] - 0.5 * m->dx[i + 1] - 0.5 * m->dx[i];
    nc = (s->Ny % 2 == 0) ? s->Ny / 2 + m->AY : (s->Ny - 1) / 2 + m->AY;
    m->Y[nc] = (s->Ny % 2 == 0) ? 0.5 * m->dy[nc] + s->y0 : s->y0;
    for (j = nc + 1; j < m->M; j++) m->Y[j] = m->Y[j - 1] + 0.5 * m->dy[j - 1] + 0.5 * m->dy[j];
    for (j = nc - 1; j >= 0; j--) m->Y[j] = m->Y[j + 1] - 0.5 * m->dy[j + 1] - 0.5 * m->dy[j];
    m->Z[0] = 0.5 * m->dz[0] + s->z0;
    for (k = 1; k < m->N; k++) m->Z[k] = m->Z[k - 1] + 0.5 * m->dz[k - 1] + 0.5 * m->dz[k];
    m->AnLx = (int)ceil((m->X[m->AL - 1] - m->X[0] + m->dx[m->AL - 1] / 2 + m->dx[0] / 2) / s->dxmin);
    m->AnRx = (int)ceil((m->X[m->L - 1] - m->X[m->L - m->AR] + m->dx[m->L - 1] / 2 +
                         m->dx[m->L - m->AR] / 2) / s->dxmin);
    m->Any = (int)ceil((m->Y[m->AY - 1] - m->Y[0] + m->dy[m->AY - 1] / 2 + m->dy[0] / 2) / s->dymin);
    m->L1 = m->AnLx + m->AnRx + s->Nx;
    m->M1 = 2 * m->Any + s->Ny;
  }
  orc_build_mapping(m);
  size_t ncell = (size_t)m->L * m->M * m->N;
  m->cond = malloc(ncell * sizeof(double));
  if (cond) memcpy(m->cond, cond, ncell * sizeof(double));
  else for (size_t c = 0; c < ncell; c++) m->cond[c] = 1.0;
  m->Ex = calloc(SZ_EX(m), sizeof(double));
  m->Ey = calloc(SZ_EY(m), sizeof(double));
  m->Ez = calloc(SZ_EZ(m), sizeof(double));
  m->Bx = calloc(SZ_BX(m), sizeof(double));
  m->By = calloc(SZ_BY(m), sizeof(double));
  m->Bz = calloc(SZ_BZ(m), sizeof(double));
  m->BxAir = calloc((size_t)m->M * (m->L + 1), sizeof(double));
  m->ByAir = calloc((size_t)(m->M + 1) * m->L, sizeof(double));
  size_t np = (size_t)m->M1 * m->L1;
  m->T = calloc((size_t)m->M * m->L1, sizeof(double));
  m->BzAir = calloc(np, sizeof(double));
  m->Fr = calloc(np, sizeof(double));
  m->Fi = calloc(np, sizeof(double));
  m->Gr = calloc(np, sizeof(double));
  m->Gi = calloc(np, sizeof(double));
  return m;
}

void orc_destroy(orc_model *m) {
  if (!m) return;
  double *p[] = {m->dx, m->dy, m->dz, m->X, m->Y, m->Z, m->Xn, m->Yn, m->cond, m->Ex, m->Ey,
                 m->Ez, m->Bx, m->By, m->Bz, m->BxAir, m->ByAir, m->T, m->BzAir, m->Fr, m->Fi,
                 m->Gr, m->Gi};
  for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
  free(m->xN2o); free(m->yN2o); free(m->xo2N); free(m->yo2N); free(m->xBx2N); free(m->yBy2N);
  free(m);
}

/* sizes: L M N AnLx AnRx Any L1 M1 */
void orc_info(const orc_model *m, int *out) {
  out[0] = m->L; out[1] = m->M; out[2] = m->N; out[3] = m->AnLx; out[4] = m->AnRx;
  out[5] = m->Any; out[6] = m->L1; out[7] = m->M1;
}

void orc_grid(const orc_model *m, double *dx, double *dy, double *dz, double *X, double *Y,
              double *Z, double *Xn, double *Yn) {
  if (dx) memcpy(dx, m->dx, m->L * sizeof(double));
  if (dy) memcpy(dy, m->dy, m->M * sizeof(double));
  if (dz) memcpy(dz, m->dz, m->N * sizeof(double));
  if (X) memcpy(X, m->X, m->L * sizeof(double));
  if (Y) memcpy(Y, m->Y, m->M * sizeof(double));
  if (Z) memcpy(Z, m->Z, m->N * sizeof(double));
  if (Xn) memcpy(Xn, m->Xn, m->L1 * sizeof(double));
  if (Yn) memcpy(Yn, m->Yn, m->M1 * sizeof(double));
}

void orc_maps(const orc_model *m, int *xN2o, int *yN2o, int *xo2N, int *yo2N, int *xBx2N,
              int *yBy2N) {
  if (xN2o) memcpy(xN2o, m->xN2o, m->L1 * sizeof(int));
  if (yN2o) memcpy(yN2o, m->yN2o, m->M1 * sizeof(int));
  if (xo2N) memcpy(xo2N, m->xo2N, m->L * sizeof(int));
  if (yo2N) memcpy(yo2N, m->yo2N, m->M * sizeof(int));
  if (xBx2N) memcpy(xBx2N, m->xBx2N, (m->L + 1) * sizeof(int));
  if (yBy2N) memcpy(yBy2N, m->yBy2N, (m->M + 1) * sizeof(int));
}

static double *orc_field_ptr(orc_model *m, int comp, size_t *n) {
  switch (comp) {
    case TCSEM_EX: *n = SZ_EX(m); return m->Ex;
    case TCSEM_EY: *n = SZ_EY(m); return m->Ey;
    case TCSEM_EZ: *n = SZ_EZ(m); return m->Ez;
    case TCSEM_BX: *n = SZ_BX(m); return m->Bx;
    case TCSEM_BY: *n = SZ_BY(m); return m->By;
    case TCSEM_BZ: *n = SZ_BZ(m); return m->Bz;
    case TCSEM_BXAIR: *n = (size_t)m->M * (m->L + 1); return m->BxAir;
    case TCSEM_BYAIR: *n = (size_t)(m->M + 1) * m->L; return m->ByAir;
  }
  *n = 0;
  return NULL;
}

int orc_get_field(orc_model *m, int comp, double *out) {
  size_t n;
  double *p = orc_field_ptr(m, comp, &n);
  if (!p) return TCSEM_EINVAL;
  memcpy(out, p, n * sizeof(double));
  return 0;
}

int orc_set_field(orc_model *m, int comp, const double *in) {
  size_t n;
  double *p = orc_field_ptr(m, comp, &n);
  if (!p) return TCSEM_EINVAL;
  memcpy(p, in, n * sizeof(double));
  return 0;
}

void orc_zero_fields(orc_model *m) { /* prepare_free_mem.cu:164-186 */
  memset(m->Ex, 0, SZ_EX(m) * sizeof(double));
  memset(m->Ey, 0, SZ_EY(m) * sizeof(double));
  memset(m->Ez, 0, SZ_EZ(m) * sizeof(double));
  memset(m->Bx, 0, SZ_BX(m) * sizeof(double));
  memset(m->By, 0, SZ_BY(m) * sizeof(double));
  memset(m->Bz, 0, SZ_BZ(m) * sizeof(double));
  memset(m->BxAir, 0, (size_t)m->M * (m->L + 1) * sizeof(double));
  memset(m->ByAir, 0, (size_t)(m->M + 1) * m->L * sizeof(double));
}

/* ============================================================================================
 * source schedule: SourceWave.cpp:8-58 (dt[2*ns] is read before it is written in the
 * reference; its malloc'd value was 0 when testwaveform.dat was produced, so 0 is used).
 * ========================================================================================== */
int orc_waveform(int ns, double sita_t, double *t, double *dt, double *wave) {
  double t0, t1;
  int i;
  if (ns < 1 || !(sita_t > 0)) return TCSEM_EINVAL;
  for (i = 0; i < 2 * ns + 2; i++) dt[i] = 0.0;
  t[0] = 0;
  for (i = 0; i < ns; i++) {
    t0 = t[i] - 3 * sita_t;
    dt[i] = -sita_t * sita_t * (1 - exp(-4.5)) / ns / t0 / exp(-t0 * t0 / (2 * sita_t * sita_t));
    t[i + 1] = t[i] + dt[i];
  }
  if (3 * sita_t < t[ns]) t[ns] = 3 * sita_t;
  for (i = ns + 1; i < 2 * ns + 1; i++) {
    t[i] = 6 * sita_t - t[2 * ns - i];
    dt[i - 1] = t[i] - t[i - 1];
  }
  for (i = 0; i < 2 * ns + 1; i++)
    wave[i] = 1 / (sqrt(2 * ORC_PI) * sita_t) *
              exp(-(t[i] - 3 * sita_t) * (t[i] - 3 * sita_t) / 2 / sita_t / sita_t);
  dt[2 * ns + 1] = 0.5 * (dt[2 * ns] * dt[2 * ns] / dt[2 * ns - 1] + dt[2 * ns]);
  t0 = dt[0];
  dt[0] = 0.5 * (t0 * t0 / dt[1] + t0);
  for (i = 1; i < 2 * ns + 1; i++) {
    t1 = dt[i];
    dt[i] = 0.5 * (t0 + dt[i]);
    t0 = t1;
  }
  return 0;
}

/* logspace.cpp:5-64 */
static void orc_logspace(double *a, double b, double e, int n) {
  double step = (log10(e) - log10(b)) / (n - 1);
  for (int i = 0; i < n; i++) a[i] = pow(10.0, log10(b) + i * step);
}
static void orc_linspace(double *a, double b, double e, int n) {
  double step = (e - b) / (n - 1);
  for (int i = 0; i < n; i++) a[i] = b + i * step;
}
int orc_search_insert(const double *A, int n, double target) {
  if (n < 1) return 0;
  int low = 0, high = n - 1, mid, pos = 0;
  while (low <= high) {
    mid = low + (high - low) / 2;
    if (A[mid] == target) return mid;
    else if (A[mid] > target) { high = mid - 1; pos = mid; }
    else { low = mid + 1; pos = mid + 1; }
  }
  return pos;
}
void orc_times(const tcsem_rxtime *rt, double *times) { /* LoadRxTxPos.cpp:438-447 */
  if (rt->is_log) orc_logspace(times, rt->T0, rt->Tn, rt->Nt);
  else orc_linspace(times, rt->T0, rt->Tn, rt->Nt);
}

/* ============================================================================================
 * up-continuation: SplineGridConvIR2R.cu:70-512 + initialize_UpCont_gpu_resources.cu:46-144
 * (GPU form, NPad=4), UpContinuation_gpu.cu:42-164, C2R semantics of cuFFT/FFTW.
 * ========================================================================================== */

/* natural spline on nodes xs (n) with spacing h_i = 0.5*d[i]+0.5*d[i+1], values ys; evaluate at
 * the new points given by (xnew[p], idraw[p] = index of the first node >= xnew[p], relative to
 * the window).  Tridiagonal solve without pivoting (cusparse gtsv2_nopivot). */
static void orc_spline_window(int n, const double *d, const double *xs, const double *ys,
                              int np, const double *xnew, const int *idraw, double *out,
                              double *b, double *cp) {
  int i;
  /* A (fillA*) and rhs (Fill_B_*) */
  double *dl = cp, *dd = cp + n, *du = cp + 2 * n, *rhs = cp + 3 * n;
  for (i = 0; i < n; i++) {
    if (i == 0 || i == n - 1) {
      dd[i] = 2.0; dl[i] = 0.0; du[i] = 0.0; rhs[i] = 0.0;
    } else {
      dd[i] = 2.0 / 3.0 * (d[i] + 0.5 * d[i - 1] + 0.5 * d[i + 1]);
      dl[i] = 1.0 / 3.0 * (0.5 * d[i - 1] + 0.5 * d[i]);
      du[i] = 1.0 / 3.0 * (0.5 * d[i] + 0.5 * d[i + 1]);
      rhs[i] = (ys[i + 1] - ys[i]) / (0.5 * d[i] + 0.5 * d[i + 1]) -
               (ys[i] - ys[i - 1]) / (0.5 * d[i - 1] + 0.5 * d[i]);
    }
  }
  /* Thomas */
  double *c2 = b; /* reuse b as modified upper diag, then overwrite with solution */
  double *r2 = cp + 4 * n;
  c2[0] = du[0] / dd[0];
  r2[0] = rhs[0] / dd[0];
  for (i = 1; i < n; i++) {
    double den = dd[i] - dl[i] * c2[i - 1];
    c2[i] = du[i] / den;
    r2[i] = (rhs[i] - dl[i] * r2[i - 1]) / den;
  }
  b[n - 1] = r2[n - 1];
  for (i = n - 2; i >= 0; i--) b[i] = r2[i] - c2[i] * b[i + 1];
  /* Intep_BzIR2R_*: coefficient per target, cubic inside, linear beyond the ends */
  for (int p = 0; p < np; p++) {
    int id = imax(idraw[p] - 1, 0);
    double ma, mc;
    if (id < n - 1) {
      double h = 0.5 * d[id] + 0.5 * d[id + 1];
      ma = 1.0 / 3.0 * (b[id + 1] - b[id]) / h;
      mc = (ys[id + 1] - ys[id]) / h - 1.0 / 3.0 * (2.0 * b[id] + b[id + 1]) * h;
    } else {
      double h = 0.5 * d[id - 1] + 0.5 * d[id];
      ma = 1.0 / 3.0 * (b[id] - b[id - 1]) / h;
      mc = (ys[id] - ys[id - 1]) / h - 1.0 / 3.0 * (2.0 * b[id - 1] + b[id]) * h;
      mc += 3.0 * ma * h * h + 2.0 * b[id - 1] * h;
      ma = 0.0;
    }
    double hh = xnew[p] - xs[id];
    int ir = idraw[p] - 1;
    if (ir >= 0 && ir < n - 1) out[p] = ma * hh * hh * hh + b[ir] * hh * hh + mc * hh + ys[ir];
    else out[p] = (b[imax(ir, 0)] * hh + mc) * hh + ys[imax(ir, 0)];
  }
}

/* Bz0 (M x L) -> BzAir (M1 x L1) */
void orc_spline_ir2r(orc_model *m, const double *bz0, double *bzair) {
  int L = m->L, M = m->M, L1 = m->L1, M1 = m->M1, Nx = m->s.Nx, Ny = m->s.Ny;
  int n1 = m->AL + ORC_NPAD, n2 = m->AR + ORC_NPAD, n3 = m->AY + ORC_NPAD;
  int ox = Nx + m->AL - ORC_NPAD, oy = Ny + m->AY - ORC_NPAD;
  double *T = m->T; /* M x L1: x pass */
#pragma omp parallel
  {
    int nmax = imax(imax(n1, n2), n3);
    double *b = malloc(nmax * sizeof(double)), *cp = malloc(5 * nmax * sizeof(double));
    int *idr = malloc(imax(L1, M1) * sizeof(int));
    double *col = malloc((M > M1 ? M : M1) * sizeof(double));
    double *res = malloc(imax(L1, M1) * sizeof(double));
#pragma omp for
    for (int j = 0; j < M; j++) {
      const double *row = bz0 + (size_t)j * L;
      double *trow = T + (size_t)j * L1;
      for (int p = 0; p < m->AnLx; p++) idr[p] = m->xN2o[p];
      orc_spline_window(n1, m->dx, m->X, row, m->AnLx, m->Xn, idr, trow, b, cp);
      for (int p = 0; p < Nx; p++) trow[m->AnLx + p] = row[m->AL + p];
      for (int p = 0; p < m->AnRx; p++) idr[p] = m->xN2o[m->AnLx + Nx + p] - ox;
      orc_spline_window(n2, m->dx + ox, m->X + ox, row + ox, m->AnRx, m->Xn + m->AnLx + Nx, idr,
                        trow + m->AnLx + Nx, b, cp);
    }
#pragma omp for
    for (int c = 0; c < L1; c++) {
      for (int j = 0; j < M; j++) col[j] = T[(size_t)j * L1 + c];
      for (int p = 0; p < m->Any; p++) idr[p] = m->yN2o[p];
      orc_spline_window(n3, m->dy, m->Y, col, m->Any, m->Yn, idr, res, b, cp);
      for (int p = 0; p < m->Any; p++) bzair[(size_t)p * L1 + c] = res[p];
      for (int p = 0; p < Ny; p++) bzair[(size_t)(m->Any + p) * L1 + c] = col[m->AY + p];
      for (int p = 0; p < m->Any; p++) idr[p] = m->yN2o[m->Any + Ny + p] - oy;
      orc_spline_window(n3, m->dy + oy, m->Y + oy, col + oy, m->Any, m->Yn + m->Any + Ny, idr,
                        res, b, cp);
      for (int p = 0; p < m->Any; p++) bzair[(size_t)(m->Any + Ny + p) * L1 + c] = res[p];
    }
    free(b); free(cp); free(idr); free(col); free(res);
  }
}

/* -- complex FFT (radix-2 iterative; Bluestein for other lengths), sign -1 forward -------- */
static void fft_pow2(int n, double *re, double *im, int sign) {
  int i, j, k, len;
  for (i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (len = 2; len <= n; len <<= 1) {
    int h = len >> 1;
    for (k = 0; k < h; k++) {
      double ang = sign * 2.0 * ORC_PI * k / len;
      double wr = cos(ang), wi = sin(ang);
      for (i = k; i < n; i += len) {
        double xr = re[i + h] * wr - im[i + h] * wi;
        double xi = re[i + h] * wi + im[i + h] * wr;
        re[i + h] = re[i] - xr; im[i + h] = im[i] - xi;
        re[i] += xr; im[i] += xi;
      }
    }
  }
}

static void fft_any(int n, double *re, double *im, int sign) {
  if ((n & (n - 1)) == 0) { fft_pow2(n, re, im, sign); return; }
  int mlen = 1;
  while (mlen < 2 * n - 1) mlen <<= 1;
  double *ar = calloc(mlen, sizeof(double)), *ai = calloc(mlen, sizeof(double));
  double *br = calloc(mlen, sizeof(double)), *bi = calloc(mlen, sizeof(double));
  double *wr = malloc(n * sizeof(double)), *wi = malloc(n * sizeof(double));
  for (int k = 0; k < n; k++) { /* w_k = exp(sign*i*pi*k^2/n), k^2 mod 2n for accuracy */
    long long q = ((long long)k * k) % (2LL * n);
    double ang = sign * ORC_PI * (double)q / n;
    wr[k] = cos(ang); wi[k] = sin(ang);
  }
  for (int k = 0; k < n; k++) { /* a = x * w */
    ar[k] = re[k] * wr[k] - im[k] * wi[k];
    ai[k] = re[k] * wi[k] + im[k] * wr[k];
  }
  br[0] = wr[0]; bi[0] = -wi[0];
  for (int k = 1; k < n; k++) {
    br[k] = br[mlen - k] = wr[k];
    bi[k] = bi[mlen - k] = -wi[k];
  }
  fft_pow2(mlen, ar, ai, -1);
  fft_pow2(mlen, br, bi, -1);
  for (int k = 0; k < mlen; k++) {
    double r = ar[k] * br[k] - ai[k] * bi[k], i = ar[k] * bi[k] + ai[k] * br[k];
    ar[k] = r; ai[k] = i;
  }
  fft_pow2(mlen, ar, ai, +1);
  for (int k = 0; k < n; k++) {
    double r = ar[k] / mlen, i = ai[k] / mlen;
    re[k] = r * wr[k] - i * wi[k];
    im[k] = r * wi[k] + i * wr[k];
  }
  free(ar); free(ai); free(br); free(bi); free(wr); free(wi);
}

/* unnormalised 2-D complex DFT of (rows x cols) arrays, in place */
void orc_fft2(int rows, int cols, double *re, double *im, int sign) {
#pragma omp parallel
  {
    double *cr = malloc(rows * sizeof(double)), *ci = malloc(rows * sizeof(double));
#pragma omp for
    for (int r = 0; r < rows; r++) fft_any(cols, re + (size_t)r * cols, im + (size_t)r * cols, sign);
#pragma omp for
    for (int c = 0; c < cols; c++) {
      for (int r = 0; r < rows; r++) { cr[r] = re[(size_t)r * cols + c]; ci[r] = im[(size_t)r * cols + c]; }
      fft_any(rows, cr, ci, sign);
      for (int r = 0; r < rows; r++) { re[(size_t)r * cols + c] = cr[r]; im[(size_t)r * cols + c] = ci[r]; }
    }
    free(cr); free(ci);
  }
}

/* FBz_to_FBxFBy (UpContinuation_gpu.cu:113-164) on the half spectrum (M1 x (L1/2+1)), then the
 * C2R of the Hermitian completion: Re of the inverse 2-D DFT (unnormalised).  Input bzair
 * (M1 x L1 real); outputs bxr, byr (M1 x L1). */
void orc_upcont_filter(orc_model *m, const double *bzair, double *bxr, double *byr) {
  int L = m->L1, M = m->M1, H = L / 2 + 1;
  size_t np = (size_t)L * M;
  double *Fr = m->Fr, *Fi = m->Fi, *Gr = m->Gr, *Gi = m->Gi;
  for (size_t p = 0; p < np; p++) { Fr[p] = bzair[p]; Fi[p] = 0.0; }
  orc_fft2(M, L, Fr, Fi, -1); /* D2Z: unnormalised forward */
  const double dzdx = m->s.dzmin / m->s.dxmin, dzdy = m->s.dzmin / m->s.dymin;
  const double rxy = m->s.dxmin / m->s.dymin, ryx = m->s.dymin / m->s.dxmin;
  /* half-spectrum filter, Hermitian completion into the full spectrum */
  for (int pass = 0; pass < 2; pass++) {
    double *out = pass == 0 ? bxr : byr;
    for (int iy = 0; iy < M; iy++) {
      for (int ix = 0; ix < H; ix++) {
        double v1 = (iy < M / 2 + 1) ? (1.0 * ORC_PI * iy) / M * dzdy
                                     : -(1.0 * ORC_PI * (M - iy)) / M * dzdy;
        double u1 = (1.0 * ORC_PI * ix) / L * dzdx;
        size_t p = (size_t)iy * L + ix;
        double ar = -Fi[p], ai = Fr[p]; /* iHilbert: i * FBz */
        double hr, hi;
        if (ix == 0 && iy == 0) {
          hr = ar; hi = ai;
        } else {
          double rr = u1 * u1 + v1 * v1;
          double upcc = exp(-sqrt(rr)) * (1.0 / sqrt(rr));
          double g = pass == 0 ? rxy * u1 * upcc : ryx * v1 * upcc;
          hr = g * ar; hi = g * ai;
        }
        Gr[p] = hr; Gi[p] = hi;
      }
    }
    for (int iy = 0; iy < M; iy++)
      for (int ix = H; ix < L; ix++) {
        size_t p = (size_t)iy * L + ix, q = (size_t)((M - iy) % M) * L + (L - ix);
        Gr[p] = Gr[q]; Gi[p] = -Gi[q];
      }
    orc_fft2(M, L, Gr, Gi, +1);
    for (size_t p = 0; p < np; p++) out[p] = Gr[p];
  }
}

/* Blinear_BxR2IR / Blinear_ByR2IR (UpContinuation_gpu.cu:42-110) */
void orc_bilinear(orc_model *m, const double *bxr, const double *byr, double *bxair,
                  double *byair) {
  int L0 = m->L, M0 = m->M, L1 = m->L1, M1 = m->M1;
  const double *X = m->X, *Y = m->Y, *Xn = m->Xn, *Yn = m->Yn;
  for (int iy = 0; iy < M0; iy++)
    for (int ix = 0; ix < L0 + 1; ix++) {
      int II = imin(imax(m->xBx2N[ix], 1), L1 - 1), JJ = imin(imax(m->yo2N[iy], 1), M1 - 1);
      double Xp = (ix == L0) ? X[L0 - 1] + 0.5 * m->dx[L0 - 1] : X[ix] - 0.5 * m->dx[ix];
      double l1 = (Xp - Xn[II - 1]) / (Xn[II] - Xn[II - 1]);
      double a0 = (1.0 - l1) * bxr[(JJ - 1) * L1 + II - 1] + l1 * bxr[(JJ - 1) * L1 + II];
      double a1 = (1.0 - l1) * bxr[JJ * L1 + II - 1] + l1 * bxr[JJ * L1 + II];
      l1 = (Y[iy] - Yn[JJ - 1]) / (Yn[JJ] - Yn[JJ - 1]);
      bxair[(size_t)iy * (L0 + 1) + ix] = ((1.0 - l1) * a0 + l1 * a1) / (L1 * M1);
    }
  for (int iy = 0; iy < M0 + 1; iy++)
    for (int ix = 0; ix < L0; ix++) {
      int II = imin(imax(m->xo2N[ix], 1), L1 - 1), JJ = imin(imax(m->yBy2N[iy], 1), M1 - 1);
      double Yp = (iy == M0) ? Y[M0 - 1] + 0.5 * m->dy[M0 - 1] : Y[iy] - 0.5 * m->dy[iy];
      double l1 = (X[ix] - Xn[II - 1]) / (Xn[II] - Xn[II - 1]);
      double a0 = (1.0 - l1) * byr[(JJ - 1) * L1 + II - 1] + l1 * byr[(JJ - 1) * L1 + II];
      double a1 = (1.0 - l1) * byr[JJ * L1 + II - 1] + l1 * byr[JJ * L1 + II];
      l1 = (Yp - Yn[JJ - 1]) / (Yn[JJ] - Yn[JJ - 1]);
      byair[(size_t)iy * L0 + ix] = ((1.0 - l1) * a0 + l1 * a1) / (L1 * M1);
    }
}

/* UpContinuation_gpu (UpContinuation_gpu.cu:166-274) on the plane bz0 (M x L). */
void orc_upcont_plane(orc_model *m, const double *bz0, double *bxair, double *byair) {
  size_t np = (size_t)m->L1 * m->M1;
  double *bxr = malloc(np * sizeof(double)), *byr = malloc(np * sizeof(double));
  orc_spline_ir2r(m, bz0, m->BzAir);
  orc_upcont_filter(m, m->BzAir, bxr, byr);
  orc_bilinear(m, bxr, byr, bxair, byair);
  free(bxr); free(byr);
}

/* ============================================================================================
 * CPU twins of the air boundary (the reference's CPU formulation; parity spec, compared with the
 * GPU formulation above per function): SplineGridConvIR2R_cpu.cpp:7-691, the filter of
 * UpContinuation.cpp:70-97 and BilinearGridConvR2IR.cpp:8-352.
 * ========================================================================================== */

/* tk::spline coefficients of SplineGridConvIR2R_cpu on n nodes ys with the CPU spacing d[i]
 * between nodes i and i+1 (the cell width dx_i, not the centre distance): natural ends (b = 0),
 * tridiagonal system solved by elimination without pivoting (tk::band_matrix::lu_solve,
 * Spline.h:221-231, for a 1-1 band); a, b, c of f_i(h) = a h^3 + b h^2 + c h + y_i, with the
 * right end continued as f_{n-1}(h) = b h^2 + c h + y_{n-1}, c_{n-1} = f'_{n-2}(x_{n-1}).
 * SplineGridConvIR2R_cpu.cpp:52-98. */
static void cpu_tk_spline(int n, const double *d, const double *ys, double *a, double *b, double *c,
                          double *work) {
  double *dl = work, *dd = work + n, *du = work + 2 * n, *r = work + 3 * n;
  for (int i = 0; i < n; i++) {
    if (i == 0 || i == n - 1) {
      dl[i] = 0.0; dd[i] = 2.0; du[i] = 0.0; r[i] = 0.0;
    } else {
      dl[i] = 1.0 / 3.0 * d[i - 1];
      dd[i] = 2.0 / 3.0 * (d[i] + d[i - 1]);
      du[i] = 1.0 / 3.0 * d[i];
      r[i] = (ys[i + 1] - ys[i]) / d[i] - (ys[i] - ys[i - 1]) / d[i - 1];
    }
  }
  for (int i = 1; i < n; i++) { /* forward elimination */
    double f = dl[i] / dd[i - 1];
    dd[i] -= f * du[i - 1];
    r[i] -= f * r[i - 1];
  }
  b[n - 1] = r[n - 1] / dd[n - 1];
  for (int i = n - 2; i >= 0; i--) b[i] = (r[i] - du[i] * b[i + 1]) / dd[i];
  for (int i = 0; i < n - 1; i++) {
    a[i] = 1.0 / 3.0 * (b[i + 1] - b[i]) / d[i];
    c[i] = (ys[i + 1] - ys[i]) / d[i] - 1.0 / 3.0 * (2.0 * b[i] + b[i + 1]) * d[i];
  }
  double h = d[n - 2];
  a[n - 1] = 0.0;
  c[n - 1] = 3.0 * a[n - 2] * h * h + 2.0 * b[n - 2] * h + c[n - 2];
}

/* Evaluation of SplineGridConvIR2R_cpu at a target: map = the new->old search index of the
 * target, off = global index of the window's node 0, xo = old node coordinates (global), xt =
 * the target's coordinate.  The offset h is taken from old node max(map - 1, 0) (global), the
 * branch from id = map - 1 - off; `last` is the index the reference tests for the beyond-the-end
 * branch (n - 1, except n1 - 1 in the right-back x pass, SplineGridConvIR2R_cpu.cpp:627). */
static double cpu_tk_eval(int n, int last, const double *a, const double *b, const double *c,
                          const double *ys, int map, int off, const double *xo, double xt) {
  double h = xt - xo[imax(map - 1, 0)];
  int id = map - 1 - off;
  if (id < 0) return (b[0] * h + c[0]) * h + ys[0];
  if (id == last) return (b[n - 1] * h + c[n - 1]) * h + ys[id];
  return ((a[id] * h + b[id]) * h + c[id]) * h + ys[id];
}

/* SplineGridConvIR2R_cpu (SplineGridConvIR2R_cpu.cpp:7-691): Bz0 (M x L) -> BzAir (M1 x L1).
 * x passes over n1 = AL + 2 / n2 = AR + 2 nodes (the padding plus two core nodes) and y passes
 * over n3 = AY + 2 nodes; the corner blocks go through a y pass of x-pass results whose two core
 * nodes are the first / last two x-pass rows -- on the right, taken two columns to the left of
 * the target column, as the reference does (:579-581, :616-618). */
void orc_spline_ir2r_cpu(orc_model *m, const double *bz0, double *bzair) {
  const int L = m->L, L1 = m->L1, Nx = m->s.Nx, Ny = m->s.Ny;
  const int xo = m->AL, xr = m->AR, yo = m->AY; /* irxL_old, irxR_old, iry_old */
  const int xn = m->AnLx, xrn = m->AnRx, yn = m->Any;
  const int n1 = xo + 2, n2 = xr + 2, n3 = yo + 2;
  const double *X = m->X, *Y = m->Y, *Xn = m->Xn, *Yn = m->Yn, *dx = m->dx, *dy = m->dy;
  const int *xm = m->xN2o, *ym = m->yN2o;
  int nmax = imax(imax(n1, n2), n3);
  double *a = malloc(nmax * sizeof(double)), *b = malloc(nmax * sizeof(double));
  double *c = malloc(nmax * sizeof(double)), *w = malloc(4 * nmax * sizeof(double));
  double *ys = malloc(nmax * sizeof(double));
  double *semi = malloc((size_t)imax(imax(xn, xrn), 1) * n3 * sizeof(double));
#define F_OLD(j, i) bz0[(size_t)(j) * L + (i)]
#define F_NEW(j, i) bzair[(size_t)(j) * L1 + (i)]
#define SEMI(i, j) semi[(size_t)(i) * n3 + (j)]
  /* core copy (:43-50) */
  for (int j = 0; j < Ny; j++)
    for (int i = 0; i < Nx; i++) F_NEW(j + yn, i + xn) = F_OLD(j + yo, i + xo);

  /* one x pass of old row j over the window starting at old column x0 (nodes x0 .. x0+nn-1),
   * evaluated at new columns c0 .. c0+nc-1; results go to dst[k] */
#define X_PASS(j, x0, nn, c0, nc, last, STORE)                                               \
  do {                                                                                        \
    for (int q = 0; q < (nn); q++) ys[q] = F_OLD(j, (x0) + q);                                \
    cpu_tk_spline((nn), dx + (x0), ys, a, b, c, w);                                           \
    for (int k = 0; k < (nc); k++) {                                                          \
      double v = cpu_tk_eval((nn), (last), a, b, c, ys, xm[(c0) + k], (x0), X, Xn[(c0) + k]); \
      STORE;                                                                                  \
    }                                                                                         \
  } while (0)
  /* one y pass over nodes ys (already filled; window starting at old row y0), evaluated at new
   * rows r0 .. r0+nr-1 */
#define Y_PASS(y0, r0, nr, STORE)                                                             \
  do {                                                                                        \
    cpu_tk_spline(n3, dy + (y0), ys, a, b, c, w);                                             \
    for (int k = 0; k < (nr); k++) {                                                          \
      double v = cpu_tk_eval(n3, n3 - 1, a, b, c, ys, ym[(r0) + k], (y0), Y, Yn[(r0) + k]);   \
      STORE;                                                                                  \
    }                                                                                         \
  } while (0)

  const int yb = yo + Ny - 2; /* old row of the back windows' node 0 */
  const int xb = Nx + xo - 2; /* old column of the right windows' node 0 */
  /* ---- left: x passes (front rows into semi, middle rows into the plane), y passes of the
   * corners (:52-346) ---- */
  for (int j = 0; j < yo; j++) X_PASS(j, 0, n1, 0, xn, n1 - 1, SEMI(k, j) = v);
  for (int j = yo; j < yo + Ny; j++) X_PASS(j, 0, n1, 0, xn, n1 - 1, F_NEW(j - yo + yn, k) = v);
  for (int j = yo; j < yo + 2; j++)
    for (int i = 0; i < xn; i++) SEMI(i, j) = F_NEW(j - yo + yn, i);
  for (int i = 0; i < xn; i++) {
    for (int q = 0; q < n3; q++) ys[q] = SEMI(i, q);
    Y_PASS(0, 0, yn, F_NEW(k, i) = v);
  }
  for (int j = yo + Ny - 2; j < yo + Ny; j++)
    for (int i = 0; i < xn; i++) SEMI(i, j + 2 - yo - Ny) = F_NEW(j - yo + yn, i);
  for (int j = yo + Ny; j < Ny + 2 * yo; j++) X_PASS(j, 0, n1, 0, xn, n1 - 1, SEMI(k, j + 2 - yo - Ny) = v);
  for (int i = 0; i < xn; i++) {
    for (int q = 0; q < n3; q++) ys[q] = SEMI(i, q);
    Y_PASS(yb, yn + Ny, yn, F_NEW(yn + Ny + k, i) = v);
  }
  /* ---- centre columns: front and back y passes (:348-432) ---- */
  for (int jc = xo; jc < xo + Nx; jc++) {
    for (int q = 0; q < n3; q++) ys[q] = F_OLD(q, jc);
    Y_PASS(0, 0, yn, F_NEW(k, jc - xo + xn) = v);
    for (int q = 0; q < n3; q++) ys[q] = F_OLD(yb + q, jc);
    Y_PASS(yb, yn + Ny, yn, F_NEW(yn + Ny + k, jc - xo + xn) = v);
  }
  /* ---- right (:435-688) ---- */
  const int cr = Nx + xn; /* first new column of the right padding */
  for (int j = 0; j < yo; j++) X_PASS(j, xb, n2, cr, xrn, n2 - 1, SEMI(k, j) = v);
  for (int j = yo; j < yo + Ny; j++) X_PASS(j, xb, n2, cr, xrn, n2 - 1, F_NEW(j - yo + yn, cr + k) = v);
  for (int j = yo; j < yo + 2; j++)
    for (int i = 0; i < xrn; i++) SEMI(i, j) = F_NEW(j - yo + yn, i + cr - 2);
  for (int i = 0; i < xrn; i++) {
    for (int q = 0; q < n3; q++) ys[q] = SEMI(i, q);
    Y_PASS(0, 0, yn, F_NEW(k, i + cr) = v);
  }
  for (int j = yo + Ny - 2; j < yo + Ny; j++)
    for (int i = 0; i < xrn; i++) SEMI(i, j + 2 - yo - Ny) = F_NEW(j - yo + yn, i + cr - 2);
  for (int j = yo + Ny; j < Ny + 2 * yo; j++) X_PASS(j, xb, n2, cr, xrn, n1 - 1, SEMI(k, j + 2 - yo - Ny) = v);
  for (int i = 0; i < xrn; i++) {
    for (int q = 0; q < n3; q++) ys[q] = SEMI(i, q);
    Y_PASS(yb, yn + Ny, yn, F_NEW(yn + Ny + k, i + cr) = v);
  }
#undef X_PASS
#undef Y_PASS
#undef F_OLD
#undef F_NEW
#undef SEMI
  free(a); free(b); free(c); free(w); free(ys); free(semi);
}

/* The filter of UpContinuation.cpp:70-97 on the half spectrum: physical wavenumbers
 * u = 2 pi ix / (L1 dxmin), v = +-2 pi iy / (M1 dymin), upcc = exp(-dzmin/2 |k|) / |k|,
 * FBx = i u upcc FBz, FBy = i v upcc FBz, DC FBx = FBy = i FBz; then the C2R of the FFTW plans
 * (Hermitian completion, real part), unnormalised.  Identical to the GPU filter when dxmin =
 * dymin; otherwise the GPU form carries the extra factors dx/dy and dy/dx
 * (UpContinuation_gpu.cu:159-160). */
void orc_upcont_filter_cpu(orc_model *m, const double *bzair, double *bxr, double *byr) {
  int L = m->L1, M = m->M1, H = L / 2 + 1;
  size_t np = (size_t)L * M;
  double *Fr = m->Fr, *Fi = m->Fi, *Gr = m->Gr, *Gi = m->Gi;
  for (size_t p = 0; p < np; p++) { Fr[p] = bzair[p]; Fi[p] = 0.0; }
  orc_fft2(M, L, Fr, Fi, -1);
  const double dz = m->s.dzmin, dxm = m->s.dxmin, dym = m->s.dymin;
  for (int pass = 0; pass < 2; pass++) {
    double *out = pass == 0 ? bxr : byr;
    for (int j = 0; j < M; j++)
      for (int i = 0; i < H; i++) {
        double v = (j < M / 2 + 1) ? 2 * ORC_PI * j / M / dym : -2 * ORC_PI * (M - j) / M / dym;
        double u = 2 * ORC_PI * i / L / dxm;
        size_t p = (size_t)j * L + i;
        double g = 1.0;
        if (!(i == 0 && j == 0)) {
          double k = sqrt(u * u + v * v);
          g = (pass == 0 ? u : v) * exp(-dz / 2 * k) / k;
        }
        Gr[p] = -g * Fi[p];
        Gi[p] = g * Fr[p];
      }
    for (int j = 0; j < M; j++)
      for (int i = H; i < L; i++) {
        size_t p = (size_t)j * L + i, q = (size_t)((M - j) % M) * L + (L - i);
        Gr[p] = Gr[q]; Gi[p] = -Gi[q];
      }
    /* self-conjugate bins: the C2R keeps their real part */
    for (int j = 0; j < M; j += (M % 2 == 0 ? M / 2 : M))
      for (int i = 0; i < L; i += (L % 2 == 0 ? L / 2 : L)) Gi[(size_t)j * L + i] = 0.0;
    orc_fft2(M, L, Gr, Gi, +1);
    for (size_t p = 0; p < np; p++) out[p] = Gr[p];
  }
}

/* BilinearGridConvR2IR (BilinearGridConvR2IR.cpp:8-352): linear interpolation between the
 * regular nodes around each Bx / By point, linear extrapolation from the first / last two nodes
 * where the search index is 0 or L1 (M1); divided by L1 M1 as UpContinuation.cpp:142-150 does.
 * Where the reference reads Bnew[JJ][L1] (one past the row, :99 and :257, the II == L1, JJ == 0
 * corner), the last-two-nodes formula of the neighbouring branches is used. */
static double cpu_bilin_point(const double *B, int L1, int M1, const double *Xn, const double *Yn,
                              int II, int JJ, double x, double y) {
  int i0, j0; /* the two nodes used in x: i0, i0 + 1; in y: j0, j0 + 1 */
  if (II == 0) i0 = 0;
  else if (II == L1) i0 = L1 - 2;
  else i0 = II - 1;
  if (JJ == 0) j0 = 0;
  else if (JJ == M1) j0 = M1 - 2;
  else j0 = JJ - 1;
  double hx = Xn[i0 + 1] - Xn[i0];
  double k1 = (B[(size_t)j0 * L1 + i0 + 1] - B[(size_t)j0 * L1 + i0]) / hx;
  double k2 = (B[(size_t)(j0 + 1) * L1 + i0 + 1] - B[(size_t)(j0 + 1) * L1 + i0]) / hx;
  double a1 = B[(size_t)j0 * L1 + i0] + k1 * (x - Xn[i0]);
  double a2 = B[(size_t)(j0 + 1) * L1 + i0] + k2 * (x - Xn[i0]);
  return a1 + (a2 - a1) / (Yn[j0 + 1] - Yn[j0]) * (y - Yn[j0]);
}

void orc_bilinear_cpu(orc_model *m, const double *bxr, const double *byr, double *bxair, double *byair) {
  int L0 = m->L, M0 = m->M, L1 = m->L1, M1 = m->M1;
  double norm = (double)L1 * M1;
  for (int j = 0; j < M0; j++)
    for (int i = 0; i < L0 + 1; i++) {
      double x = (i == L0) ? m->X[L0 - 1] + 0.5 * m->dx[L0 - 1] : m->X[i] - 0.5 * m->dx[i];
      bxair[(size_t)j * (L0 + 1) + i] =
          cpu_bilin_point(bxr, L1, M1, m->Xn, m->Yn, m->xBx2N[i], m->yo2N[j], x, m->Y[j]) / norm;
    }
  for (int j = 0; j < M0 + 1; j++)
    for (int i = 0; i < L0; i++) {
      double y = (j == M0) ? m->Y[M0 - 1] + 0.5 * m->dy[M0 - 1] : m->Y[j] - 0.5 * m->dy[j];
      byair[(size_t)j * L0 + i] =
          cpu_bilin_point(byr, L1, M1, m->Xn, m->Yn, m->xo2N[i], m->yBy2N[j], m->X[i], y) / norm;
    }
}

/* UpContinuation.cpp:8-211: the CPU-twin chain on the plane bz0 (M x L). */
void orc_upcont_plane_cpu(orc_model *m, const double *bz0, double *bxair, double *byair) {
  size_t np = (size_t)m->L1 * m->M1;
  double *bxr = malloc(np * sizeof(double)), *byr = malloc(np * sizeof(double));
  orc_spline_ir2r_cpu(m, bz0, m->BzAir);
  orc_upcont_filter_cpu(m, m->BzAir, bxr, byr);
  orc_bilinear_cpu(m, bxr, byr, bxair, byair);
  free(bxr); free(byr);
}

void orc_set_cpu_twins(orc_model *m, int on) { m->cpu_twins = on != 0; }

void orc_upcont(orc_model *m) {
  if (m->cpu_twins) orc_upcont_plane_cpu(m, m->Bz, m->BxAir, m->ByAir);
  else orc_upcont_plane(m, m->Bz, m->BxAir, m->ByAir);
}

/* ============================================================================================
 * E update (ElectricalFiled.cu:89-345), race-free semantics.
 * ========================================================================================== */
void orc_update_e(orc_model *m, double lambda, double dt) {
  const int L = m->L, M = m->M, N = m->N;
  const double *c = m->cond, *dx = m->dx, *dy = m->dy, *dz = m->dz;
  double *Ex = m->Ex, *Ey = m->Ey, *Ez = m->Ez;
  const double *Bx = m->Bx, *By = m->By, *Bz = m->Bz;
  const double miu = ORC_MIU;
#define CN(k, j, i) c[((size_t)(k) * M + (j)) * L + (i)]
  /* UpdateEx: i in [0,L), j in [1,M), k in [0,N) */
#pragma omp parallel for
  for (int j = 1; j < M; j++) {
    for (int i = 0; i < L; i++) {
      const double dy2 = dy[j - 1] + dy[j];
      double sx = 0.5 * (CN(0, j - 1, i) + CN(0, j, i)), sy = 0.0;
      double ave = 0.5 * (sx + sy);
      double term1 = (Bz[(size_t)j * L + i] - Bz[(size_t)(j - 1) * L + i]) / dy2;
      double byx = By[(size_t)j * L + i], byy = m->ByAir[(size_t)j * L + i];
      size_t e = (size_t)j * L + i;
      Ex[e] = (2.0 * lambda - ave * dt) / (2.0 * lambda + ave * dt) * Ex[e] +
              4.0 * dt / (2.0 * lambda + ave * dt) * (term1 - (byx - byy) / (2.0 * dz[0])) / miu;
      for (int k = 1; k < N; k++) {
        sy = sx;
        sx = 0.5 * (CN(k, j - 1, i) + CN(k, j, i));
        ave = 0.5 * (sx + sy);
        term1 = (Bz[((size_t)k * M + j) * L + i] - Bz[((size_t)k * M + j - 1) * L + i]) / dy2;
        byy = byx;
        byx = By[((size_t)k * (M + 1) + j) * L + i];
        e = ((size_t)k * (M + 1) + j) * L + i;
        Ex[e] = (2.0 * lambda - ave * dt) / (2.0 * lambda + ave * dt) * Ex[e] +
                4.0 * dt / (2.0 * lambda + ave * dt) * (term1 - (byx - byy) / (dz[k] + dz[k - 1])) /
                    miu;
      }
    }
  }
  /* UpdateEy: i in [1,L), j in [0,M) */
#pragma omp parallel for
  for (int j = 0; j < M; j++) {
    for (int i = 1; i < L; i++) {
      const double dx2 = dx[i - 1] + dx[i];
      double sita = 0.25 * (CN(0, j, i - 1) + CN(0, j, i));
      double bxx = Bx[(size_t)j * (L + 1) + i], bxy = m->BxAir[(size_t)j * (L + 1) + i];
      size_t e = (size_t)j * (L + 1) + i;
      Ey[e] = (2.0 * lambda - sita * dt) / (2.0 * lambda + sita * dt) * Ey[e] +
              4.0 * dt / (2.0 * lambda + sita * dt) *
                  ((bxx - bxy) / (2.0 * dz[0]) - (Bz[(size_t)j * L + i] - Bz[(size_t)j * L + i - 1]) / dx2) /
                  miu;
      for (int k = 1; k < N; k++) {
        sita = 0.25 * (CN(k, j, i - 1) + CN(k, j, i) + CN(k - 1, j, i - 1) + CN(k - 1, j, i));
        bxy = bxx;
        bxx = Bx[((size_t)k * M + j) * (L + 1) + i];
        e = ((size_t)k * M + j) * (L + 1) + i;
        size_t bz = ((size_t)k * M + j) * L + i;
        Ey[e] = (2.0 * lambda - sita * dt) / (2.0 * lambda + sita * dt) * Ey[e] +
                4.0 * dt / (2.0 * lambda + sita * dt) *
                    ((bxx - bxy) / (dz[k] + dz[k - 1]) - (Bz[bz] - Bz[bz - 1]) / dx2) / miu;
      }
    }
  }
  /* UpdateEz: i in [1,L), j in [1,M), k in [0,N) */
#pragma omp parallel for
  for (int j = 1; j < M; j++) {
    for (int i = 1; i < L; i++) {
      const double dx2 = dx[i - 1] + dx[i], dy2 = dy[j - 1] + dy[j];
      for (int k = 0; k < N; k++) {
        double sita = 0.25 * (CN(k, j - 1, i - 1) + CN(k, j - 1, i) + CN(k, j, i - 1) + CN(k, j, i));
        double term2 = Bx[((size_t)k * M + j) * (L + 1) + i] - Bx[((size_t)k * M + j - 1) * (L + 1) + i];
        size_t by = ((size_t)k * (M + 1) + j) * L + i;
        size_t e = ((size_t)k * (M + 1) + j) * (L + 1) + i;
        Ez[e] = (2.0 * lambda - sita * dt) / (2.0 * lambda + sita * dt) * Ez[e] +
                4.0 * dt / (2.0 * lambda + sita * dt) * ((By[by] - By[by - 1]) / dx2 - term2 / dy2) / miu;
      }
    }
  }
#undef CN
}

/* ============================================================================================
 * transmitters: LoadRxTxPos.cpp:143-428 (cell search), ElectricalFiled.cu:347-641 (injection)
 * ========================================================================================== */
typedef struct orc_txpos {
  int ort; double BipoLen, x, y, z, azi, dip;
  int ixa, iya, iza, ixb, iyb, izb, ictx, icty, ictz;
  double BipoLen1, x1, y1, z1, azi1, dip1;
  int ixa1, iya1, iza1, ixb1, iyb1, izb1, ictx1, icty1, ictz1;
} orc_txpos;

static void orc_pair(int *a, int *b) { /* LoadRxTxPos.cpp:297-302 */
  if (*b > *a) (*b)--;
  else if (*b < *a) (*a)--;
}

/* returns TCSEM_EDOMAIN for a transmitter outside the model (the reference exits) */
int orc_locate_tx(const orc_model *m, const tcsem_tx *t, orc_txpos *p) {
  const int L = m->L, M = m->M, K = m->N;
  memset(p, 0, sizeof(*p));
  p->ort = t->ort; p->BipoLen = t->bipolen; p->x = t->x; p->y = t->y; p->z = t->z;
  p->BipoLen1 = t->bipolen1; p->x1 = t->x1; p->y1 = t->y1; p->z1 = t->z1;
  if (t->ort == 4 || t->ort == 5) {
    p->azi = t->azi_deg * ORC_PI / 180.0; p->dip = t->dip_deg * ORC_PI / 180.0;
  }
  if (t->ort == 5) {
    p->azi1 = t->azi1_deg * ORC_PI / 180.0; p->dip1 = t->dip1_deg * ORC_PI / 180.0;
  }
  switch (t->ort) {
    case 1:
      p->ixa = orc_search_insert(m->X, L, p->x - p->BipoLen / 2);
      p->ixb = orc_search_insert(m->X, L, p->x + p->BipoLen / 2) - 1;
      if (p->ixa == 0 || p->ixb == L) return TCSEM_EDOMAIN;
      p->icty = orc_search_insert(m->Y, M, p->y);
      if (p->icty == 0 || p->icty == M) return TCSEM_EDOMAIN;
      p->ictz = orc_search_insert(m->Z, K, p->z);
      break;
    case 2:
      p->ictx = orc_search_insert(m->X, L, p->x);
      if (p->ictx == 0 || p->ictx == L) return TCSEM_EDOMAIN;
      p->iya = orc_search_insert(m->Y, M, p->y - p->BipoLen / 2);
      p->iyb = orc_search_insert(m->Y, M, p->y + p->BipoLen / 2) - 1;
      if (p->iya == 0 || p->iyb == M) return TCSEM_EDOMAIN;
      p->ictz = orc_search_insert(m->Z, K, p->z);
      break;
    case 3:
      p->ictx = orc_search_insert(m->X, L, p->x);
      if (p->ictx == 0 || p->ictx == L) return TCSEM_EDOMAIN;
      p->icty = orc_search_insert(m->Y, M, p->y);
      if (p->icty == 0 || p->icty == M) return TCSEM_EDOMAIN;
      p->iza = orc_search_insert(m->Z, K, p->z - p->BipoLen / 2);
      p->izb = orc_search_insert(m->Z, K, p->z + p->BipoLen / 2) - 1;
      break;
    case 4:
    case 5: {
      double cx = p->BipoLen * cos(p->dip) * cos(p->azi) / 2.0;
      double cy = p->BipoLen * cos(p->dip) * sin(p->azi) / 2.0;
      double cz = p->BipoLen * sin(p->dip) / 2.0;
      p->ixa = orc_search_insert(m->X, L, p->x - cx);
      p->ixb = orc_search_insert(m->X, L, p->x + cx);
      orc_pair(&p->ixa, &p->ixb);
      p->iya = orc_search_insert(m->Y, M, p->y - cy);
      p->iyb = orc_search_insert(m->Y, M, p->y + cy);
      orc_pair(&p->iya, &p->iyb);
      p->iza = orc_search_insert(m->Z, K, p->z - cz);
      p->izb = orc_search_insert(m->Z, K, p->z + cz);
      orc_pair(&p->iza, &p->izb);
      p->ictx = orc_search_insert(m->X, L, p->x);
      p->icty = orc_search_insert(m->Y, M, p->y);
      p->ictz = orc_search_insert(m->Z, K, p->z);
      if (t->ort == 5) {
        double c1x = p->BipoLen1 * cos(p->dip1) * cos(p->azi1) / 2.0;
        double c1y = p->BipoLen1 * cos(p->dip1) * sin(p->azi1) / 2.0;
        double c1z = p->BipoLen1 * sin(p->dip1) / 2.0;
        p->ixa1 = orc_search_insert(m->X, L, p->x1 - c1x);
        p->ixb1 = orc_search_insert(m->X, L, p->x1 + c1x);
        orc_pair(&p->ixa1, &p->ixb1);
        p->iya1 = orc_search_insert(m->Y, M, p->y1 - c1y);
        p->iyb1 = orc_search_insert(m->Y, M, p->y1 + c1y);
        orc_pair(&p->iya1, &p->iyb1);
        p->iza1 = orc_search_insert(m->Z, K, p->z1 - c1z);
        p->izb1 = orc_search_insert(m->Z, K, p->z1 + c1z);
        orc_pair(&p->iza1, &p->izb1);
        p->ictx1 = orc_search_insert(m->X, L, p->x1);
        p->icty1 = orc_search_insert(m->Y, M, p->y1);
        p->ictz1 = orc_search_insert(m->Z, K, p->z1);
        if (p->ixa1 == 0 || p->ixb1 == L || p->iya1 == 0 || p->iyb1 == M) return TCSEM_EDOMAIN;
      }
      if (p->ixa == 0 || p->ixb == L || p->iya == 0 || p->iyb == M) return TCSEM_EDOMAIN;
      break;
    }
    default:
      return TCSEM_EINVAL;
  }
  return 0;
}

/* AddTx_x/y/z (ElectricalFiled.cu:347-475); nthreads mirrors the launch width. */
static void orc_addtx_x(orc_model *m, const orc_txpos *t, double cur, int nthreads) {
  const int L = m->L, M = m->M;
  const int ixa = imin(t->ixa, t->ixb), iya = t->icty, iza = t->ictz;
  const double *c = m->cond;
  for (int ix = 0; ix < nthreads; ix++) {
    if (ix > abs(t->ixb - t->ixa)) continue;
    double sita;
    if (iza == 0) sita = c[((size_t)iza * M + iya) * L + ix + ixa] + c[((size_t)iza * M + iya - 1) * L + ix + ixa];
    else
      sita = c[((size_t)iza * M + iya) * L + ix + ixa] + c[((size_t)iza * M + iya - 1) * L + ix + ixa] +
             c[((size_t)(iza - 1) * M + iya) * L + ix + ixa] + c[((size_t)(iza - 1) * M + iya - 1) * L + ix + ixa];
    m->Ex[((size_t)iza * (M + 1) + iya) * L + ix + ixa] -= 2.0 * cur / (0.25 * sita * m->dy[iya] * m->dz[iza]);
  }
}
static void orc_addtx_y(orc_model *m, const orc_txpos *t, double cur, int nthreads) {
  const int L = m->L, M = m->M;
  const int ixa = t->ictx, iya = imin(t->iya, t->iyb), iza = t->ictz;
  const double *c = m->cond;
  for (int iy = 0; iy < nthreads; iy++) {
    if (iy > abs(t->iyb - t->iya)) continue;
    double sita;
    if (iza == 0) sita = c[((size_t)iza * M + iy + iya) * L + ixa - 1] + c[((size_t)iza * M + iy + iya) * L + ixa];
    else
      sita = c[((size_t)iza * M + iy + iya) * L + ixa - 1] + c[((size_t)iza * M + iy + iya) * L + ixa] +
             c[((size_t)(iza - 1) * M + iy + iya) * L + ixa - 1] + c[((size_t)(iza - 1) * M + iy + iya) * L + ixa];
    m->Ey[((size_t)iza * M + iy + iya) * (L + 1) + ixa] -= 2.0 * cur / (0.25 * sita * m->dx[ixa] * m->dz[iza]);
  }
}
static void orc_addtx_z(orc_model *m, const orc_txpos *t, double cur, int nthreads) {
  const int L = m->L, M = m->M;
  const int ixa = t->ictx, iya = t->icty, iza = imin(t->iza, t->izb);
  const double *c = m->cond;
  for (int iz = 0; iz < nthreads; iz++) {
    if (iz > abs(t->izb - t->iza)) continue;
    double sita = c[((size_t)(iza + iz) * M + iya) * L + ixa - 1] + c[((size_t)(iza + iz) * M + iya) * L + ixa] +
                  c[((size_t)(iza + iz) * M + iya - 1) * L + ixa - 1] + c[((size_t)(iza + iz) * M + iya - 1) * L + ixa];
    m->Ez[((size_t)(iza + iz) * (M + 1) + iya) * (L + 1) + ixa] -= 2.0 * cur / (0.25 * sita * m->dx[ixa] * m->dy[iya]);
  }
}

/* host side of ElectricalField (ElectricalFiled.cu:532-641), including its use of the first
 * bipole's cells for both bipoles of ort 5. */
void orc_add_source(orc_model *m, const orc_txpos *t, double wave) {
  if (wave == 0) return;
  const double *X = m->X, *Y = m->Y, *Z = m->Z, *dx = m->dx, *dy = m->dy, *dz = m->dz;
  double cur;
  switch (t->ort) {
    case 1:
      wave /= (X[t->ixb] - X[t->ixa] + dx[t->ixb] / 2 + dx[t->ixa] / 2);
      orc_addtx_x(m, t, wave, abs(t->ixb - t->ixa) + 1);
      break;
    case 2:
      wave /= (Y[t->iyb] - Y[t->iya] + dy[t->iyb] / 2 + dy[t->iya] / 2);
      orc_addtx_y(m, t, wave, abs(t->iyb - t->iya) + 1);
      break;
    case 3:
      wave /= (Z[t->izb] - Z[t->iza] + dz[t->izb] / 2 + dz[t->iza] / 2);
      orc_addtx_z(m, t, wave, abs(t->izb - t->iza) + 1);
      break;
    case 4:
    case 5: {
      double f = t->ort == 5 ? t->BipoLen : 1.0;
      for (int b = 0; b < (t->ort == 5 ? 2 : 1); b++) {
        double dip = b ? t->dip1 : t->dip, azi = b ? t->azi1 : t->azi;
        int ixa = b ? t->ixa1 : t->ixa, ixb = b ? t->ixb1 : t->ixb;
        int iya = b ? t->iya1 : t->iya, iyb = b ? t->iyb1 : t->iyb;
        int iza = b ? t->iza1 : t->iza, izb = b ? t->izb1 : t->izb;
        if (b) f = t->BipoLen1;
        if (t->ort == 5)
          cur = wave * cos(dip) * cos(azi) * f / (fabs(X[ixb] - X[ixa]) + dx[ixb] / 2.0 + dx[ixa] / 2.0);
        else
          cur = wave * cos(dip) * cos(azi) / (fabs(X[ixb] - X[ixa]) + dx[ixb] / 2.0 + dx[ixa] / 2.0);
        orc_addtx_x(m, t, cur, abs(ixb - ixa) + 1);
        if (t->ort == 5)
          cur = wave * cos(dip) * sin(azi) * f / (fabs(Y[iyb] - Y[iya]) + dy[iyb] / 2.0 + dy[iya] / 2.0);
        else
          cur = wave * cos(dip) * sin(azi) / (fabs(Y[iyb] - Y[iya]) + dy[iyb] / 2.0 + dy[iya] / 2.0);
        orc_addtx_y(m, t, cur, abs(iyb - iya) + 1);
        if (t->ort == 5)
          cur = wave * sin(dip) * f / (fabs(Z[izb] - Z[iza]) + dz[izb] / 2.0 + dz[iza] / 2.0);
        else
          cur = wave * sin(dip) / (fabs(Z[izb] - Z[iza]) + dz[izb] / 2.0 + dz[iza] / 2.0);
        orc_addtx_z(m, t, cur, abs(izb - iza) + 1);
      }
      break;
    }
  }
}

/* ============================================================================================
 * B update (Bderivate.cu:79-265)
 * ========================================================================================== */
void orc_update_b(orc_model *m, double dt) {
  const int L = m->L, M = m->M, N = m->N;
  const double *dx = m->dx, *dy = m->dy, *dz = m->dz;
  const double *Ex = m->Ex, *Ey = m->Ey, *Ez = m->Ez;
  double *Bx = m->Bx, *By = m->By, *Bz = m->Bz;
#pragma omp parallel for
  for (int k = 0; k < N; k++) {
    for (int j = 0; j < M; j++)
      for (int i = 1; i < L; i++) {
        double ey0 = Ey[((size_t)k * M + j) * (L + 1) + i], ey1 = Ey[((size_t)(k + 1) * M + j) * (L + 1) + i];
        double ez0 = Ez[((size_t)k * (M + 1) + j) * (L + 1) + i], ez1 = Ez[((size_t)k * (M + 1) + j + 1) * (L + 1) + i];
        Bx[((size_t)k * M + j) * (L + 1) + i] += ((ey1 - ey0) / dz[k] - (ez1 - ez0) / dy[j]) * dt;
      }
    for (int j = 1; j < M; j++)
      for (int i = 0; i < L; i++) {
        double ez0 = Ez[((size_t)k * (M + 1) + j) * (L + 1) + i], ez1 = Ez[((size_t)k * (M + 1) + j) * (L + 1) + i + 1];
        double ex0 = Ex[((size_t)k * (M + 1) + j) * L + i], ex1 = Ex[((size_t)(k + 1) * (M + 1) + j) * L + i];
        By[((size_t)k * (M + 1) + j) * L + i] += ((ez1 - ez0) / dx[i] - (ex1 - ex0) / dz[k]) * dt;
      }
  }
#pragma omp parallel for
  for (int j = 0; j < M; j++)
    for (int i = 0; i < L; i++) {
      double bz0 = 0.0;
      for (int k = N - 1; k >= 0; k--) {
        size_t bx = ((size_t)k * M + j) * (L + 1) + i, by = ((size_t)k * (M + 1) + j) * L + i;
        bz0 += dz[k] * ((Bx[bx + 1] - Bx[bx]) / dx[i] + (By[by + L] - By[by]) / dy[j]);
        Bz[((size_t)k * M + j) * L + i] = bz0;
      }
    }
}

/* Bz of the current Bx, By (UpdateBz alone), e.g. to make a caller-set state consistent. */
void orc_update_bz(orc_model *m) {
  const int L = m->L, M = m->M, N = m->N;
#pragma omp parallel for
  for (int j = 0; j < M; j++)
    for (int i = 0; i < L; i++) {
      double bz0 = 0.0;
      for (int k = N - 1; k >= 0; k--) {
        size_t bx = ((size_t)k * M + j) * (L + 1) + i, by = ((size_t)k * (M + 1) + j) * L + i;
        bz0 += m->dz[k] * ((m->Bx[bx + 1] - m->Bx[bx]) / m->dx[i] + (m->By[by + L] - m->By[by]) / m->dy[j]);
        m->Bz[((size_t)k * M + j) * L + i] = bz0;
      }
    }
}

/* one full DF step as in main.cu:223-248: up-continuation, E + source, B */
void orc_step(orc_model *m, double lambda, double dt1, double wave, const orc_txpos *tx,
              double dtB) {
  orc_upcont(m);
  orc_update_e(m, lambda, dt1);
  if (tx) orc_add_source(m, tx, wave);
  orc_update_b(m, dtB);
}

/* ============================================================================================
 * receivers: EFInterp2Rx.cu:46-470
 * ========================================================================================== */
static double spline4(const double *x, const double *y, double t) { /* CubicSpline_1d4p */
  const double h1 = x[2] - x[1];
  const double v1 = 2.0 * (x[2] - x[0]);
  const double v2 = 2.0 * (x[3] - x[1]);
  const double u1 = 6.0 * ((y[2] - y[1]) / h1 - (y[1] - y[0]) / (x[1] - x[0]));
  const double u2 = 6.0 * ((y[3] - y[2]) / (x[3] - x[2]) - (y[2] - y[1]) / h1);
  const double z1 = (v2 * u1 - h1 * u2) / (v1 * v2 - h1 * h1);
  const double z2 = (v1 * u2 - h1 * u1) / (v1 * v2 - h1 * h1);
  const double a = t - x[1], b = x[2] - t;
  return z2 / (6.0 * h1) * (a * a * a) + z1 / (6.0 * h1) * (b * b * b) + (y[2] / h1 - z2 / 6.0 * h1) * a +
         (y[1] / h1 - z1 / 6.0 * h1) * b;
}
double orc_spline4(const double *x, const double *y, double t) { return spline4(x, y, t); }

typedef struct orc_rx {
  int nrx;
  double *x, *y, *z;
  int *ix, *iy, *iz;
} orc_rx;

/* Stencil of component comp for a receiver at (x, y, z) whose searchInsert indices are (ix, iy,
 * iz) (EFInterp2Rx.cu:75-88 and the same choice in every *SpatialInterp kernel): the first node
 * index s[0..2] of the 4 nodes per axis and the node coordinates t[0..3] (x), t[4..7] (y),
 * t[8..11] (z).  Centre axes start at i-2; edge axes (nodes at the cell's lower face) use the
 * half-cell test; the z start is clamped to level 0 (the reference reads level -1 there). */
static void orc_stencil(const orc_model *m, double x, double y, double z, int ix, int iy, int iz, int comp,
                        int *s, double *t) {
  const double *X = m->X, *Y = m->Y, *Z = m->Z, *dx = m->dx, *dy = m->dy, *dz = m->dz;
  /* centre (c) or edge (e) stencils per axis; edge axes use the half-cell test */
  int xe = 0, ye = 0, ze = 0;
  switch (comp) {
    case 0: xe = 0; ye = 1; ze = 1; break; /* Ex: x-centre, y-edge, z-edge */
    case 1: xe = 1; ye = 0; ze = 1; break; /* Ey */
    case 2: xe = 1; ye = 1; ze = 0; break; /* Ez */
    case 3: xe = 1; ye = 0; ze = 0; break; /* Bx */
    case 4: xe = 0; ye = 1; ze = 0; break; /* By */
    case 5: xe = 0; ye = 0; ze = 1; break; /* Bz */
  }
  int s1 = xe ? ((X[ix - 1] + 0.5 * dx[ix - 1]) > x ? ix - 2 : ix - 1) : ix - 2;
  int s2 = ye ? ((Y[iy - 1] + 0.5 * dy[iy - 1]) > y ? iy - 2 : iy - 1) : iy - 2;
  int s3;
  if (ze) s3 = (iz >= 1 && (Z[iz - 1] + 0.5 * dz[iz - 1]) > z) ? iz - 2 : iz - 1;
  else s3 = iz - 2;
  if (s3 < 0) s3 = 0;
  s[0] = s1; s[1] = s2; s[2] = s3;
  for (int q = 0; q < 4; q++) {
    t[q] = xe ? X[s1 + q] - 0.5 * dx[s1 + q] : X[s1 + q];
    t[4 + q] = ye ? Y[s2 + q] - 0.5 * dy[s2 + q] : Y[s2 + q];
    t[8 + q] = ze ? Z[s3 + q] - 0.5 * dz[s3 + q] : Z[s3 + q];
  }
}

/* stencil of one receiver position, indices searched as LoadRxTxPos.cpp:93-121 does */
void orc_rx_stencil(const orc_model *m, double x, double y, double z, int comp, int *s, double *t) {
  const int ix = orc_search_insert(m->X, m->L, x), iy = orc_search_insert(m->Y, m->M, y);
  const int iz = orc_search_insert(m->Z, m->N, z);
  orc_stencil(m, x, y, z, ix, iy, iz, comp, s, t);
}

/* component c of receiver r: spatial interpolation from the current fields */
static double orc_rx_value(const orc_model *m, const orc_rx *R, int r, int comp) {
  const int L = m->L, M = m->M;
  const double x = R->x[r], y = R->y[r], z = R->z[r];
  int s[3];
  double t[12];
  orc_stencil(m, x, y, z, R->ix[r], R->iy[r], R->iz[r], comp, s, t);
  const int s1 = s[0], s2 = s[1], s3 = s[2];
  const double *tx = t, *ty = t + 4, *tz = t + 8;
  const double *F;
  int pitch, rows;
  switch (comp) {
    case 0: F = m->Ex; pitch = L; rows = M + 1; break;
    case 1: F = m->Ey; pitch = L + 1; rows = M; break;
    case 2: F = m->Ez; pitch = L + 1; rows = M + 1; break;
    case 3: F = m->Bx; pitch = L + 1; rows = M; break;
    case 4: F = m->By; pitch = L; rows = M + 1; break;
    default: F = m->Bz; pitch = L; rows = M; break;
  }
  double Y1[4], Y2[4];
  for (int ik = 0; ik < 4; ik++) {
    for (int ij = 0; ij < 4; ij++)
      Y1[ij] = spline4(tx, F + ((size_t)(ik + s3) * rows + ij + s2) * pitch + s1, x);
    Y2[ik] = spline4(ty, Y1, y);
  }
  return spline4(tz, Y2, z);
}

/* Spatial interpolation of the current fields to nrx receivers (x, y, z per receiver), for the
 * ncomp component ids comps: out[r * ncomp + c] (the gathers of EFXInterp / BFXInterp,
 * EFInterp2Rx.cu:411-470, without the time interpolation). */
void orc_rx_values(const orc_model *m, int nrx, const double *rx_xyz, int ncomp, const int *comps,
                   double *out) {
  orc_rx R;
  double x, y, z;
  int ix, iy, iz;
  R.nrx = 1; R.x = &x; R.y = &y; R.z = &z; R.ix = &ix; R.iy = &iy; R.iz = &iz;
  for (int r = 0; r < nrx; r++) {
    x = rx_xyz[3 * r]; y = rx_xyz[3 * r + 1]; z = rx_xyz[3 * r + 2];
    ix = orc_search_insert(m->X, m->L, x);
    iy = orc_search_insert(m->Y, m->M, y);
    iz = orc_search_insert(m->Z, m->N, z);
    for (int c = 0; c < ncomp; c++) out[(size_t)r * ncomp + c] = orc_rx_value(m, &R, 0, comps[c]);
  }
}

/* ============================================================================================
 * per-transmitter run: main.cu:195-298 with EFInterp2Rx.cu:472-585 (counters reset per Tx)
 * ========================================================================================== */
/* Returns steps taken or a negative error.  efs: [nr][NComp][Nt]. */
int orc_run(orc_model *m, const tcsem_rxtime *rt, int nrx, const double *rx_xyz,
            const tcsem_tx *tx, int max_steps, double *efs) {
  const tcsem_setup *s = &m->s;
  const int ns = s->n_src / 2;
  int rc;
  orc_txpos tp;
  if ((rc = orc_locate_tx(m, tx, &tp)) != 0) return rc;
  orc_rx R;
  R.nrx = nrx;
  R.x = malloc(nrx * sizeof(double)); R.y = malloc(nrx * sizeof(double)); R.z = malloc(nrx * sizeof(double));
  R.ix = malloc(nrx * sizeof(int)); R.iy = malloc(nrx * sizeof(int)); R.iz = malloc(nrx * sizeof(int));
  for (int r = 0; r < nrx; r++) {
    R.x[r] = rx_xyz[3 * r]; R.y[r] = rx_xyz[3 * r + 1]; R.z[r] = rx_xyz[3 * r + 2];
    R.ix[r] = orc_search_insert(m->X, m->L, R.x[r]);
    R.iy[r] = orc_search_insert(m->Y, m->M, R.y[r]);
    R.iz[r] = orc_search_insert(m->Z, m->N, R.z[r]);
  }
  int ncomp = 0, nefcmp = 0, comps[6];
  for (int c = 0; c < 6; c++) if (rt->is_comp[c]) comps[ncomp++] = c;
  for (int c = 0; c < 3; c++) if (rt->is_comp[c]) nefcmp++;
  const int nr = tx->nr, Nt = rt->Nt;
  double *times = malloc(Nt * sizeof(double));
  orc_times(rt, times);
  double *ef0 = calloc((size_t)2 * ncomp * nr, sizeof(double)); /* zeroed: the reference is not */
  memset(efs, 0, (size_t)nr * ncomp * Nt * sizeof(double));
  double *t = malloc((2 * ns + 1) * sizeof(double)), *dtw = malloc((2 * ns + 2) * sizeof(double));
  double *wave = malloc((2 * ns + 1) * sizeof(double));
  orc_waveform(ns, s->sigma_t, t, dtw, wave);
  const double miu = ORC_MIU, minsita = 0.1, mindx = m->mindx;
  orc_zero_fields(m);
  int n = 0, ne = 0, nb = 0;
  double Tpof[2][3] = {{0}};
  double toT = 0, dt1 = dtw[n], dt2 = dtw[n + 1];
  Tpof[0][1] = toT + dt1;
  Tpof[1][1] = toT + dt1 + 0.5 * dt2;
  int has_b = rt->is_comp[3] || rt->is_comp[4] || rt->is_comp[5];
  int steps = 0;
  while (ne < Nt || (nb < Nt && nefcmp < ncomp)) {
    if (max_steps > 0 && steps >= max_steps) break;
    double nmuda = 7.5 * s->sfa2 * dt2 * dt1 / miu / mindx / mindx;
    orc_step(m, nmuda, dt1, n < 2 * ns ? wave[n] : 0.0, &tp, 0.5 * (dt1 + dt2));
    toT += dt1;
    dt1 = dt2;
    n++;
    if (n < 2 * ns) dt2 = dtw[n + 1];
    else dt2 = fmin(0.15 * s->sfa1 * sqrt(miu * toT * minsita / 8.) * mindx, 1.2 * dt1);
    steps++;
    /* EFInterp2Rx(t0 = toT + dt1, dt = dt2) */
    double t0 = toT + dt1, l1;
    for (int fld = 0; fld < 2; fld++) {
      int *cnt = fld == 0 ? &ne : &nb;
      if (*cnt >= Nt) continue;
      if (fld == 1 && !has_b) continue;
      double *T = Tpof[fld];
      T[2] = fld == 0 ? t0 : t0 + 0.5 * dt2;
      int c0 = fld == 0 ? 0 : nefcmp, c1 = fld == 0 ? nefcmp : ncomp;
      if (times[*cnt] <= T[0])
        while (*cnt < Nt && times[*cnt] <= T[0]) (*cnt)++;
      if (*cnt < Nt && times[*cnt] <= T[2] && times[*cnt] > T[1]) {
        for (int ir = 0; ir < nr; ir++)
          for (int c = c0; c < c1; c++)
            ef0[(size_t)c * nr + ir] = orc_rx_value(m, &R, tx->rx_ids[ir] - 1, comps[c]);
      } else if (*cnt < Nt && times[*cnt] <= T[1] && times[*cnt] > T[0]) {
        for (int ir = 0; ir < nr; ir++)
          for (int c = c0; c < c1; c++)
            ef0[(size_t)ncomp * nr + (size_t)c * nr + ir] = orc_rx_value(m, &R, tx->rx_ids[ir] - 1, comps[c]);
        while (*cnt < Nt && times[*cnt] <= T[1]) {
          l1 = (times[*cnt] - T[0]) / (T[1] - T[0]);
          for (int ir = 0; ir < nr; ir++)
            for (int c = c0; c < c1; c++)
              efs[(size_t)ir * ncomp * Nt + (size_t)c * Nt + *cnt] =
                  (1.0 - l1) * ef0[(size_t)c * nr + ir] + l1 * ef0[(size_t)ncomp * nr + (size_t)c * nr + ir];
          (*cnt)++;
        }
      }
      T[0] = T[1];
      T[1] = T[2];
    }
  }
  free(times); free(ef0); free(t); free(dtw); free(wave);
  free(R.x); free(R.y); free(R.z); free(R.ix); free(R.iy); free(R.iz);
  return steps;
}

/* The field-independent step schedule of main.cu:149-273: per step n, lambda, dt1 (E), dtB (B)
 * and the source amplitude Wave[n] (0 after the source).  Returns the number of steps written
 * (<= cap) until all samples are taken, or a negative error. */
int orc_schedule(const tcsem_setup *s, double mindx, const tcsem_rxtime *rt, int cap,
                 double *lam, double *dte, double *dtb, double *amp) {
  const int ns = s->n_src / 2;
  double *t = malloc((2 * ns + 1) * sizeof(double)), *dtw = malloc((2 * ns + 2) * sizeof(double));
  double *wave = malloc((2 * ns + 1) * sizeof(double));
  orc_waveform(ns, s->sigma_t, t, dtw, wave);
  const double miu = ORC_MIU, minsita = 0.1;
  int nefcmp = 0, ncomp = 0;
  for (int c = 0; c < 6; c++) ncomp += rt->is_comp[c] != 0;
  for (int c = 0; c < 3; c++) nefcmp += rt->is_comp[c] != 0;
  int has_b = rt->is_comp[3] || rt->is_comp[4] || rt->is_comp[5];
  double *times = malloc(rt->Nt * sizeof(double));
  orc_times(rt, times);
  int n = 0, ne = 0, nb = 0, Nt = rt->Nt;
  double Tpof[2][3] = {{0}};
  double toT = 0, dt1 = dtw[0], dt2 = dtw[1];
  Tpof[0][1] = dt1;
  Tpof[1][1] = dt1 + 0.5 * dt2;
  int steps = 0;
  while (ne < Nt || (nb < Nt && nefcmp < ncomp)) {
    if (steps >= cap) break;
    lam[steps] = 7.5 * s->sfa2 * dt2 * dt1 / miu / mindx / mindx;
    dte[steps] = dt1;
    dtb[steps] = 0.5 * (dt1 + dt2);
    amp[steps] = n < 2 * ns ? wave[n] : 0.0;
    toT += dt1;
    dt1 = dt2;
    n++;
    if (n < 2 * ns) dt2 = dtw[n + 1];
    else dt2 = fmin(0.15 * s->sfa1 * sqrt(miu * toT * minsita / 8.) * mindx, 1.2 * dt1);
    steps++;
    double t0 = toT + dt1;
    for (int fld = 0; fld < 2; fld++) {
      int *cnt = fld == 0 ? &ne : &nb;
      if (*cnt >= Nt || (fld == 1 && !has_b)) continue;
      double *T = Tpof[fld];
      T[2] = fld == 0 ? t0 : t0 + 0.5 * dt2;
      while (*cnt < Nt && times[*cnt] <= T[0]) (*cnt)++;
      if (*cnt < Nt && times[*cnt] <= T[1] && times[*cnt] > T[0])
        while (*cnt < Nt && times[*cnt] <= T[1]) (*cnt)++;
      T[0] = T[1];
      T[1] = T[2];
    }
  }
  free(t); free(dtw); free(wave); free(times);
  return steps;
}

double orc_mindx(const orc_model *m) { return m->mindx; }
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
