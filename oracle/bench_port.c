/* -----------------------------------------------------------------------------
 * CPU ORACLE (test infrastructure): the N_Vector microbenchmark on the host,
 * the CPU baseline beside the device vector ops.
 *
 * nthreads == 1: the serial vector restatement (nvector_serial_port.c)
 *   through the generic N_V* entry points, the serial vector's loops
 *   (nvector_serial.c:387-480 LinearSum, :605-622 DotProd, :641-668
 *   WrmsNorm, :994-1026 DotProdMulti: one pass over x per vector).
 * nthreads > 1: the OpenMP vector's loops restated with the same pragmas
 *   (nvector_openmp.c:424-490 LinearSum -> VLin / Vaxpy, :698-716 DotProd
 *   with reduction(+), :1001-1019 WSqrSum, :1214-1262 DotProdMulti: one
 *   parallel region, one omp-for reduction per vector), schedule(static).
 * Data: x, y, Y_k ~ U(-1, 1), w ~ U(1e-3, 1), first touched by the threads
 * that use them.  Times: milliseconds per call, after one warm-up call.
 * ---------------------------------------------------------------------------*/
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

static double unif(uint64_t seed, uint64_t i, double lo, double hi)
{
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + i;
  z          = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z          = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return lo + (hi - lo) * (((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0));
}

static void fill(double* v, long n, uint64_t seed, double lo, double hi, int nt)
{
#pragma omp parallel for schedule(static) num_threads(nt)
  for (long i = 0; i < n; i++) { v[i] = unif(seed, (uint64_t)i, lo, hi); }
}

/* OpenMP forms (nthreads > 1) */
static void omp_linsum(double a, const double* x, double b, const double* y, double* z, long n,
                       int nt)
{
#pragma omp parallel for schedule(static) num_threads(nt)
  for (long i = 0; i < n; i++) { z[i] = (a * x[i]) + (b * y[i]); }
}

static void omp_axpy1(const double* x, double* y, long n, int nt)
{
#pragma omp parallel for schedule(static) num_threads(nt)
  for (long i = 0; i < n; i++) { y[i] += x[i]; }
}

static double omp_dot(const double* x, const double* y, long n, int nt)
{
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static) num_threads(nt)
  for (long i = 0; i < n; i++) { sum += x[i] * y[i]; }
  return sum;
}

static double omp_wrms(const double* x, const double* w, long n, int nt)
{
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static) num_threads(nt)
  for (long i = 0; i < n; i++)
  {
    const double p = x[i] * w[i];
    sum += p * p;
  }
  return sqrt(sum / (double)n);
}

static void omp_dotmulti(int nvec, const double* x, double* const* Y, double* d, long n, int nt)
{
  for (int k = 0; k < nvec; k++) { d[k] = 0.0; }
#pragma omp parallel num_threads(nt)
  {
    for (int k = 0; k < nvec; k++)
    {
      const double* yd = Y[k];
      double sum       = 0.0;
#pragma omp for schedule(static)
      for (long j = 0; j < n; j++) { sum += x[j] * yd[j]; }
#pragma omp critical
      {
        d[k] += sum;
      }
    }
  }
}

int oracle_bench_vector_ops(long n, int nvec, int reps, int nthreads, double* out_ms)
{
  if (n <= 0 || nvec < 1 || nvec > 32 || reps < 1 || nthreads < 1 || !out_ms) { return -1; }
  SUNContext ctx = NULL;
  if (SUNContext_Create(SUN_COMM_NULL, &ctx)) { return -1; }
  int rc          = -1;
  N_Vector v[4]   = {NULL, NULL, NULL, NULL};
  N_Vector Yv[32] = {NULL};
  double* Yd[32]  = {NULL};
  double d[32];
  const int fill_threads = nthreads > 1 ? nthreads : 1;
  for (int i = 0; i < 4; i++)
  {
    v[i] = N_VNew_Serial(n, ctx);
    if (!v[i]) { goto done; }
  }
  for (int k = 0; k < nvec; k++)
  {
    Yv[k] = N_VNew_Serial(n, ctx);
    if (!Yv[k]) { goto done; }
    Yd[k] = N_VGetArrayPointer(Yv[k]);
    fill(Yd[k], n, 10 + (uint64_t)k, -1.0, 1.0, fill_threads);
  }
  {
    N_Vector x = v[0], y = v[1], z = v[2], w = v[3];
    double *xd = N_VGetArrayPointer(x), *yd = N_VGetArrayPointer(y);
    double *zd = N_VGetArrayPointer(z), *wd = N_VGetArrayPointer(w);
    fill(xd, n, 1, -1.0, 1.0, fill_threads);
    fill(yd, n, 2, -1.0, 1.0, fill_threads);
    fill(zd, n, 4, -1.0, 1.0, fill_threads);
    fill(wd, n, 3, 1e-3, 1.0, fill_threads);
    volatile double sink = 0.0;
    for (int op = 0; op < 5; op++)
    {
      double t0 = 0.0;
      for (int r = -1; r < reps; r++)
      {
        if (r == 0) { t0 = omp_get_wtime(); }
        if (nthreads == 1)
        {
          switch (op)
          {
          case 0: N_VLinearSum(0.3, x, 0.7, y, z); break;
          case 1: sink += N_VWrmsNorm(x, w); break;
          case 2: N_VDotProdMulti(nvec, x, Yv, d); sink += d[0]; break;
          case 3: N_VLinearSum(1.0, x, 1.0, y, y); break;
          case 4: sink += N_VDotProd(x, y); break;
          }
        }
        else
        {
          switch (op)
          {
          case 0: omp_linsum(0.3, xd, 0.7, yd, zd, n, nthreads); break;
          case 1: sink += omp_wrms(xd, wd, n, nthreads); break;
          case 2: omp_dotmulti(nvec, xd, Yd, d, n, nthreads); sink += d[0]; break;
          case 3: omp_axpy1(xd, yd, n, nthreads); break;
          case 4: sink += omp_dot(xd, yd, n, nthreads); break;
          }
        }
      }
      out_ms[op] = 1e3 * (omp_get_wtime() - t0) / reps;
    }
    (void)sink;
    rc = 0;
  }
done:
  for (int i = 0; i < 4; i++)
  {
    if (v[i]) { N_VDestroy(v[i]); }
  }
  for (int k = 0; k < nvec; k++)
  {
    if (Yv[k]) { N_VDestroy(Yv[k]); }
  }
  SUNContext_Free(&ctx);
  return rc;
}
