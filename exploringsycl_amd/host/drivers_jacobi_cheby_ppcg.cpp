// Drivers of the reference's other solvers: Jacobi (control flow of
// TeaLeaf/drivers/jacobi_driver.c:7-84), Chebyshev
// (TeaLeaf/drivers/cheby_driver.c:11-183) and CG with polynomial
// preconditioning by Chebyshev inner steps, PPCG
// (TeaLeaf/drivers/ppcg_driver.c:10-189), with the eigenvalue estimate both
// Chebyshev solvers start from (TeaLeaf/drivers/eigenvalue_driver.c:11-67).
// Stop tests, print lines and the order of kernel_interface.h calls are the
// reference's, so a run gives the reference's iteration counts.
#include <float.h>
#include <math.h>

#include <vector>

#include "comms.h"
#include "drivers.h"
#include "kernel_interface.h"
#include "tea_host.h"

#define FOR_CHUNKS(settings, cc) \
  for (int cc = 0; cc < (settings)->num_chunks_per_rank; ++cc)

// CG iterations the eigenvalue estimate needs before an error-driven switch,
// and the error below which a preset switch may happen
// (TeaLeaf/shared.h:50-51).
static const int kCgItersForEigenvalues = 20;
static const double kErrorSwitchMax = 1.0;

// Records the printed counts: `printed` (Jacobi, or the CG steps of a
// Chebyshev/PPCG solve), `after` (iterations after the switch, -1 for
// Jacobi) and `estimated`; `tt` is the loop count.  The eigenvalue bounds
// exist once a solve has switched.
static void record_iterations(Chunk* chunks, Settings* settings, int tt,
                              int printed, int after, int estimated) {
  if (g_run.steps.empty()) return;
  StepRecord& s = g_run.steps.back();
  s.iterations = printed;
  s.executed = tt < settings->max_iters ? tt + 1 : tt;
  s.solver_iterations = after;
  s.estimated = estimated;
  if (after > 0) {
    s.eigmin = chunks[0].eigmin;
    s.eigmax = chunks[0].eigmax;
  }
}

// ---- Jacobi -----------------------------------------------------------------

void jacobi_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                   double* error) {
  FOR_CHUNKS(settings, cc) {
    run_jacobi_init(&chunks[cc], settings, rx, ry);
    run_copy_u(&chunks[cc], settings);
  }
  reset_fields_to_exchange(settings);
  settings->fields_to_exchange[FIELD_U] = true;

  int tt;
  for (tt = 0; tt < settings->max_iters; ++tt) {
    // The step's error is the summed update |u_new - u|; every 50th step it
    // is replaced by the squared norm of the true residual (the reference
    // backend's 2-norm entry point assigns its result).
    FOR_CHUNKS(settings, cc) run_jacobi_iterate(&chunks[cc], settings, error);
    if (tt % 50 == 0) {
      halo_update_driver(chunks, settings, 1);
      *error = 0.0;
      FOR_CHUNKS(settings, cc) {
        run_calculate_residual(&chunks[cc], settings);
        run_calculate_2norm(&chunks[cc], settings, chunks[cc].r, error);
      }
    }
    sum_over_ranks(settings, error);
    halo_update_driver(chunks, settings, 1);
    if (fabs(*error) < settings->eps) break;
  }
  print_and_log(settings, "Jacobi: \t\t%d iterations\n", tt);
  record_iterations(chunks, settings, tt, tt, -1, -1);
}

// ---- eigenvalue estimate ------------------------------------------------------

// Eigenvalues of the symmetric tridiagonal matrix with diagonal d[0..n) and
// couplings e[1..n) (e[k] joins rows k-1 and k), by the implicit QL method
// with Wilkinson's shift (Bowdler, Martin, Reinsch and Wilkinson, Numer.
// Math. 11 (1968) 293-306; the tqli form of Numerical Recipes 11.3).  d is
// overwritten by the eigenvalues.  The rotations are evaluated in the
// reference routine's operand order (TeaLeaf/drivers/eigenvalue_driver.c:
// 71-122), including what it does when a rotation radius underflows to zero
// (the sweep goes on with the next rotation), so the bounds, and with them
// every Chebyshev coefficient, are those of the reference host.
static void tridiagonal_eigenvalues(std::vector<double>& d,
                                    std::vector<double>& e) {
  const int n = (int)d.size();
  if (n == 0) return;
  // off[k] couples rows k and k+1.
  std::vector<double> off(n, 0.0);
  for (int k = 0; k + 1 < n; ++k) off[k] = e[k + 1];

  for (int l = 0; l < n; ++l) {
    int sweeps = 0;
    for (;;) {
      // First negligible coupling at or after row l splits the matrix.
      int m = l;
      while (m < n - 1) {
        const double scale = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(off[m]) + scale == scale) break;
        ++m;
      }
      if (m == l) break;
      if (sweeps++ == 30)
        die(__LINE__, __FILE__, "Too many iterations in TQLI routine\n");

      // Wilkinson shift from the leading 2x2 block, folded into g.
      double g = (d[l + 1] - d[l]) / (2.0 * off[l]);
      double rad = sqrt((g * g) + 1.0);
      const double signed_rad = g < 0 ? -fabs(rad) : fabs(rad);
      g = d[m] - d[l] + off[l] / (g + signed_rad);

      // Chase the bulge from row m up to row l with plane rotations.
      double s = 1.0, c = 1.0, shift = 0.0;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * off[i];
        const double b = c * off[i];
        rad = sqrt(f * f + g * g);
        off[i + 1] = rad;
        if (rad == 0.0) {
          d[i + 1] -= shift;
          off[m] = 0.0;
          continue;
        }
        s = f / rad;
        c = g / rad;
        g = d[i + 1] - shift;
        rad = (d[i] - g) * s + 2.0 * c * b;
        shift = s * rad;
        d[i + 1] = g + shift;
        g = c * rad - b;
      }
      d[l] = d[l] - shift;
      off[l] = g;
      off[m] = 0.0;
    }
  }
}

// Extreme eigenvalues of the CG iteration's Lanczos matrix: diagonal
// 1/alpha_k + beta_(k-1)/alpha_(k-1), off-diagonal sqrt(beta_k)/alpha_k
// (Golub and Van Loan, Matrix Computations, sec. 10.2), widened by 5% each
// way as the reference does.
static void eigenvalue_estimate(Chunk* chunks, Settings* settings,
                                int num_cg_iters) {
  FOR_CHUNKS(settings, cc) {
    Chunk& ch = chunks[cc];
    std::vector<double> diag(num_cg_iters, 0.0), off(num_cg_iters, 0.0);
    for (int k = 0; k < num_cg_iters; ++k) {
      diag[k] = 1.0 / ch.cg_alphas[k];
      if (k > 0) diag[k] += ch.cg_betas[k - 1] / ch.cg_alphas[k - 1];
      if (k < num_cg_iters - 1)
        off[k + 1] = sqrt(ch.cg_betas[k]) / ch.cg_alphas[k];
    }
    tridiagonal_eigenvalues(diag, off);

    ch.eigmin = DBL_MAX;
    ch.eigmax = DBL_MIN;
    for (int k = 0; k < num_cg_iters; ++k) {
      ch.eigmin = diag[k] < ch.eigmin ? diag[k] : ch.eigmin;
      ch.eigmax = diag[k] > ch.eigmax ? diag[k] : ch.eigmax;
    }
    if (ch.eigmin < 0.0 || ch.eigmax < 0.0)
      die(__LINE__, __FILE__, "Calculated negative eigenvalues.\n");
    ch.eigmin *= 0.95;
    ch.eigmax *= 1.05;
    print_and_log(settings,
                  "Min. eigenvalue: \t%.12e\nMax. eigenvalue: \t%.12e\n",
                  ch.eigmin, ch.eigmax);
  }
}

// Coefficients of the Chebyshev iteration on [eigmin, eigmax]: with
// theta the interval's centre, delta its half width and sigma = theta/delta,
// rho_k = T_k(sigma) / T_(k+1)(sigma) from the three-term recurrence of the
// Chebyshev polynomials, alpha_k = rho_k rho_(k-1), beta_k = 2 rho_k / delta
// (Saad, Iterative Methods for Sparse Linear Systems, alg. 12.1).
static void cheby_coefficients(Chunk* chunks, Settings* settings, int count) {
  FOR_CHUNKS(settings, cc) {
    Chunk& ch = chunks[cc];
    ch.theta = (ch.eigmax + ch.eigmin) / 2.0;
    const double delta = (ch.eigmax - ch.eigmin) / 2.0;
    const double sigma = ch.theta / delta;
    double rho_prev = 1.0 / sigma;
    for (int k = 0; k < count; ++k) {
      const double rho = 1.0 / (2.0 * sigma - rho_prev);
      ch.cheby_alphas[k] = rho * rho_prev;
      ch.cheby_betas[k] = 2.0 * rho / delta;
      rho_prev = rho;
    }
  }
}

// Whether a CG-started solve moves to its Chebyshev phase at step tt.
static bool switch_from_cg(Settings* settings, int tt, double error) {
  return settings->error_switch
             ? error < settings->eps_lim && tt > kCgItersForEigenvalues
             : tt > settings->presteps && error < kErrorSwitchMax;
}

// ---- Chebyshev ------------------------------------------------------------------

static void cheby_step(Chunk* chunks, Settings* settings, int k,
                       bool with_norm, double* error) {
  FOR_CHUNKS(settings, cc)
    run_cheby_iterate(&chunks[cc], settings, chunks[cc].cheby_alphas[k],
                      chunks[cc].cheby_betas[k]);
  if (!with_norm) return;
  *error = 0.0;
  FOR_CHUNKS(settings, cc)
    run_calculate_2norm(&chunks[cc], settings, chunks[cc].r, error);
  sum_over_ranks(settings, error);
}

void cheby_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                  double* error) {
  double rro = 0.0;
  int est_iterations = 0;
  int cheby_iters = 0;
  cg_init_driver(chunks, settings, rx, ry, &rro);

  int tt;
  for (tt = 0; tt < settings->max_iters; ++tt) {
    if (!cheby_iters && !switch_from_cg(settings, tt, *error)) {
      cg_main_step_driver(chunks, settings, tt, &rro, error);
    } else if (++cheby_iters == 1) {
      eigenvalue_estimate(chunks, settings, tt);
      cheby_coefficients(chunks, settings, settings->max_iters - tt);
      double bb = 0.0;
      FOR_CHUNKS(settings, cc) {
        run_calculate_2norm(&chunks[cc], settings, chunks[cc].u0, &bb);
        run_cheby_init(&chunks[cc], settings);
      }
      reset_fields_to_exchange(settings);
      settings->fields_to_exchange[FIELD_U] = true;
      halo_update_driver(chunks, settings, 1);
      sum_over_ranks(settings, &bb);

      cheby_step(chunks, settings, cheby_iters, true, error);

      // Steps for the Chebyshev error bound 2 gamma^k to fall from the
      // first step's residual to machine precision of |b|, with
      // gamma = (sqrt(kappa) - 1) / (sqrt(kappa) + 1); evaluated in single
      // precision as the reference does (cheby_driver.c:146-160).
      const double kappa = chunks[0].eigmax / chunks[0].eigmin;
      const double it_alpha = DBL_EPSILON * bb / (4.0 * (*error));
      const double gamma = (sqrt(kappa) - 1.0) / (sqrt(kappa) + 1.0);
      est_iterations = (int)roundf(logf(it_alpha) / (2.0 * logf(gamma)));
    } else {
      const bool with_norm = cheby_iters >= est_iterations && (tt + 1) % 10 == 0;
      cheby_step(chunks, settings, cheby_iters, with_norm, error);
    }
    halo_update_driver(chunks, settings, 1);
    if (fabs(*error) < settings->eps) break;
  }
  print_and_log(settings, "CG: \t\t\t%d iterations\n", tt - cheby_iters + 1);
  print_and_log(settings, "Cheby: \t\t\t%d iterations (%d estimated)\n",
                cheby_iters, est_iterations);
  record_iterations(chunks, settings, tt, tt - cheby_iters + 1, cheby_iters,
                    est_iterations);
}

// ---- PPCG -------------------------------------------------------------------------

// The preconditioner: ppcg_inner_steps Chebyshev steps on A sd = r started
// from sd = r / theta, accumulating into u and updating r.
static void ppcg_inner_steps(Chunk* chunks, Settings* settings) {
  FOR_CHUNKS(settings, cc) run_ppcg_init(&chunks[cc], settings);
  reset_fields_to_exchange(settings);
  settings->fields_to_exchange[FIELD_SD] = true;
  for (int k = 0; k < settings->ppcg_inner_steps; ++k) {
    halo_update_driver(chunks, settings, 1);
    FOR_CHUNKS(settings, cc)
      run_ppcg_inner_iteration(&chunks[cc], settings,
                               chunks[cc].cheby_alphas[k],
                               chunks[cc].cheby_betas[k]);
  }
  reset_fields_to_exchange(settings);
  settings->fields_to_exchange[FIELD_P] = true;
}

static void ppcg_step(Chunk* chunks, Settings* settings, double* rro,
                      double* error) {
  double pw = 0.0;
  FOR_CHUNKS(settings, cc) run_cg_calc_w(&chunks[cc], settings, &pw);
  sum_over_ranks(settings, &pw);
  const double alpha = *rro / pw;

  double rrn = 0.0;
  FOR_CHUNKS(settings, cc) run_cg_calc_ur(&chunks[cc], settings, alpha, &rrn);

  ppcg_inner_steps(chunks, settings);

  rrn = 0.0;
  FOR_CHUNKS(settings, cc)
    run_calculate_2norm(&chunks[cc], settings, chunks[cc].r, &rrn);
  sum_over_ranks(settings, &rrn);

  const double beta = rrn / *rro;
  FOR_CHUNKS(settings, cc) run_cg_calc_p(&chunks[cc], settings, beta);
  *error = rrn;
  *rro = rrn;
}

void ppcg_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                 double* error) {
  double rro = 0.0;
  int ppcg_iters = 0;
  cg_init_driver(chunks, settings, rx, ry, &rro);

  int tt;
  for (tt = 0; tt < settings->max_iters; ++tt) {
    if (!ppcg_iters && !switch_from_cg(settings, tt, *error)) {
      cg_main_step_driver(chunks, settings, tt, &rro, error);
    } else {
      if (++ppcg_iters == 1) {
        eigenvalue_estimate(chunks, settings, tt);
        cheby_coefficients(chunks, settings, settings->ppcg_inner_steps);
        FOR_CHUNKS(settings, cc) run_calculate_residual(&chunks[cc], settings);
        reset_fields_to_exchange(settings);
        settings->fields_to_exchange[FIELD_P] = true;
        halo_update_driver(chunks, settings, 1);
        sum_over_ranks(settings, &rro);
      }
      ppcg_step(chunks, settings, &rro, error);
    }
    halo_update_driver(chunks, settings, 1);
    if (fabs(*error) < settings->eps) break;
  }
  print_and_log(settings, "CG: \t\t\t%d iterations\n", tt - ppcg_iters + 1);
  print_and_log(settings,
                "PPCG: \t\t\t%d iterations (%d inner iterations per)\n",
                ppcg_iters, settings->ppcg_inner_steps);
  record_iterations(chunks, settings, tt, tt - ppcg_iters + 1, ppcg_iters, -1);
}
