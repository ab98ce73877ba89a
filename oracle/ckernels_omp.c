/*
 * CPU baseline: TeaLeaf's CG path in the style of the reference's OpenMP
 * c_kernels (one `#pragma omp parallel for` over rows per kernel, scalar
 * results by `reduction(+:)`, the unfused dataflow cg_calc_w, cg_calc_ur,
 * cg_calc_p with the depth-1 U/P halo update between iterations), built
 * with the reference's GNU flags (-O3 -march=native -funroll-loops
 * -fopenmp, TeaLeaf/make.flags:58,77).  TEST AND BENCHMARK INFRASTRUCTURE
 * ONLY: bench.py times it beside the GPU; nothing in the product links it.
 *
 * The kernels follow TeaLeaf/c_kernels/sycl/cg.cpp (cg_init, calc_w,
 * calc_ur, calc_p), solver_methods.cpp (copy_u, finalise),
 * local_halos.cpp (the reflective updates) and the CG driver's control flow
 * (TeaLeaf/drivers/cg_driver.c:7-124).  Unlike the oracle (tea_oracle.c),
 * it sums with OpenMP's reduction and lets the compiler contract products:
 * its results agree with the oracle's to rounding, not bit for bit.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define SMVP(a)                                                          \
  ((1.0 + (kx[i + 1] + kx[i]) + (ky[i + x] + ky[i])) * a[i] -            \
   (kx[i + 1] * a[i + 1] + kx[i] * a[i - 1]) - (ky[i + x] * a[i + x] + ky[i] * a[i - x]))

/* Reflective update of `depth` cells on all four faces (local_halos.cpp). */
static void halo(int x, int y, int h, int depth, double* b) {
#pragma omp parallel for
  for (int jj = h; jj < y - h; ++jj)
    for (int k = 0; k < depth; ++k) {
      b[jj * x + h - 1 - k] = b[jj * x + h + k];
      b[jj * x + x - h + k] = b[jj * x + x - h - 1 - k];
    }
#pragma omp parallel for
  for (int kk = 0; kk < x; ++kk)
    for (int k = 0; k < depth; ++k) {
      b[(y - h + k) * x + kk] = b[(y - h - 1 - k) * x + kk];
      b[(h - 1 - k) * x + kk] = b[(h + k) * x + kk];
    }
}

/* cg_init_u / _k / _others (cg.cpp:7-134): returns r.p. */
static double cg_init(int x, int y, int h, int coefficient, double rx, double ry,
                      const double* density, const double* energy, double* u, double* p,
                      double* r, double* w, double* kx, double* ky) {
#pragma omp parallel for
  for (int jj = 1; jj < y - 1; ++jj)
    for (int kk = 1; kk < x - 1; ++kk) {
      const int i = kk + jj * x;
      p[i] = 0.0;
      r[i] = 0.0;
      u[i] = energy[i] * density[i];
      w[i] = coefficient == 1 ? density[i] : 1.0 / density[i];
    }
#pragma omp parallel for
  for (int jj = h; jj < y - 1; ++jj)
    for (int kk = h; kk < x - 1; ++kk) {
      const int i = kk + jj * x;
      kx[i] = rx * (w[i - 1] + w[i]) / (2.0 * w[i - 1] * w[i]);
      ky[i] = ry * (w[i - x] + w[i]) / (2.0 * w[i - x] * w[i]);
    }
  double rro = 0.0;
#pragma omp parallel for reduction(+ : rro)
  for (int jj = h; jj < y - h; ++jj)
    for (int kk = h; kk < x - h; ++kk) {
      const int i = kk + jj * x;
      const double smvp = SMVP(u);
      w[i] = smvp;
      r[i] = u[i] - w[i];
      p[i] = r[i];
      rro += r[i] * p[i];
    }
  return rro;
}

static double calc_w(int x, int y, int h, double* w, const double* p, const double* kx,
                     const double* ky) {
  double pw = 0.0;
#pragma omp parallel for reduction(+ : pw)
  for (int jj = h; jj < y - h; ++jj)
    for (int kk = h; kk < x - h; ++kk) {
      const int i = kk + jj * x;
      const double smvp = SMVP(p);
      w[i] = smvp;
      pw += w[i] * p[i];
    }
  return pw;
}

static double calc_ur(int x, int y, int h, double alpha, double* u, double* r,
                      const double* p, const double* w) {
  double rrn = 0.0;
#pragma omp parallel for reduction(+ : rrn)
  for (int jj = h; jj < y - h; ++jj)
    for (int kk = h; kk < x - h; ++kk) {
      const int i = kk + jj * x;
      u[i] += alpha * p[i];
      r[i] -= alpha * w[i];
      rrn += r[i] * r[i];
    }
  return rrn;
}

static void calc_p(int x, int y, int h, double beta, double* p, const double* r) {
#pragma omp parallel for
  for (int jj = h; jj < y - h; ++jj)
    for (int kk = h; kk < x - h; ++kk) {
      const int i = kk + jj * x;
      p[i] = beta * p[i] + r[i];
    }
}

/*
 * `steps` timesteps of the CG path on one chunk with all faces external,
 * from the state in density and energy0 (x*y arrays, halos filled; energy0
 * is updated to the final energy after each step as finalise does).  Each
 * step: DENSITY/ENERGY halos at depth 2, cg_init, U/P halos, then CG
 * iterations until sqrt(|rr|) < eps or max_iters (cg_driver.c), at most
 * `max_total` iterations over the whole call (0: no cap).  Returns the
 * seconds spent in the solves (omp_get_wtime); iters_out[s] receives the
 * iterations each step executed (the count the driver prints, plus one on
 * convergence), *steps_out the steps run, and *temp_out the temperature
 * sum of m*u over the interior after the last step (field_summary.cpp:
 * 6-151; the cell volume dx*dy is passed in).
 */
double omp_cg_run(int x, int y, int h, int coefficient, double rx, double ry, double eps,
                  int max_iters, int steps, long max_total, double volume,
                  const double* density, double* energy, int* iters_out, int* steps_out,
                  double* temp_out) {
  const size_t n = (size_t)x * y;
  double* u = (double*)calloc(n, sizeof(double));
  double* p = (double*)calloc(n, sizeof(double));
  double* r = (double*)calloc(n, sizeof(double));
  double* w = (double*)calloc(n, sizeof(double));
  double* kx = (double*)calloc(n, sizeof(double));
  double* ky = (double*)calloc(n, sizeof(double));
  double* dens = (double*)malloc(n * sizeof(double));
  memcpy(dens, density, n * sizeof(double));
  double seconds = 0.0;
  long total = 0;
  for (int s = 0; s < steps; ++s) {
    halo(x, y, h, 2, dens);
    halo(x, y, h, 2, energy);
    const double t0 = omp_get_wtime();
    double rro = cg_init(x, y, h, coefficient, rx, ry, dens, energy, u, p, r, w, kx, ky);
    halo(x, y, h, 1, u);
    halo(x, y, h, 1, p);
    int tt = 0, executed = 0;
    for (tt = 0; tt < max_iters; ++tt) {
      if (max_total > 0 && total >= max_total) break;
      ++executed;
      const double pw = calc_w(x, y, h, w, p, kx, ky);
      const double alpha = rro / pw;
      const double rrn = calc_ur(x, y, h, alpha, u, r, p, w);
      const double beta = rrn / rro;
      calc_p(x, y, h, beta, p, r);
      halo(x, y, h, 1, u);
      halo(x, y, h, 1, p);
      rro = rrn;
      ++total;
      if (sqrt(fabs(rro)) < eps) break;
    }
    seconds += omp_get_wtime() - t0;
    iters_out[s] = executed;
    *steps_out = s + 1;
    /* finalise: energy = u / density */
#pragma omp parallel for
    for (int jj = h; jj < y - h; ++jj)
      for (int kk = h; kk < x - h; ++kk) {
        const int i = kk + jj * x;
        energy[i] = u[i] / dens[i];
      }
    if (max_total > 0 && total >= max_total) break;
  }
  double temp = 0.0;
#pragma omp parallel for reduction(+ : temp)
  for (int jj = h; jj < y - h; ++jj)
    for (int kk = h; kk < x - h; ++kk) {
      const int i = kk + jj * x;
      temp += volume * dens[i] * u[i];
    }
  *temp_out = temp;
  free(u);
  free(p);
  free(r);
  free(w);
  free(kx);
  free(ky);
  free(dens);
  return seconds;
}

int omp_threads(void) { return omp_get_max_threads(); }
