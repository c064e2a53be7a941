/*
 * edb_oracle.c -- CPU oracle (test infrastructure) for the CalQ path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference
 * algorithm, used by tests/, smoke() and bench.py's cpu_baseline leg to check
 * the HIP library's results.  It is never linked into or called by the HIP library.
 *
 * Every function cites the reference lines it restates (paths relative to
 * the reference repository).  Arithmetic is written operation by operation in
 * the reference's order and the file is compiled with -ffp-contract=off, so
 * that for odd N genB/calQ reproduce the reference bit for bit (checked
 * against golden vectors written by the compiled reference headers).
 *
 * Pinning of each function (see tests/test_oracle.py):
 *   edbo_funF, edbo_filtering, edbo_rval, edbo_genB, edbo_beta
 *       golden beta entries and checksums from QCalculator::genB (odd N,
 *       alpha 0/1, filter on/off); F_alpha against numerical quadrature of
 *       int_0^1 r^{2+alpha} sinc(xi r) sinc(eta r) dr; sigma(0) == 1.
 *   edbo_calQ, edbo_calQ_direct
 *       golden Q from QCalculator::calQ (odd N); brute-force (l,m)-order sum
 *       (even N too); Q_0 == 0; Hermitian symmetry; Q(Maxwellian) ~ 0.
 *   edbo_calQ_modes
 *       equal bit for bit to edbo_calQ with edbo_genB's B on the selected
 *       modes (odd and even N), which the checks above pin.
 *   edbo_dft3, edbo_fft, edbo_ifft      numpy.fft.fftn / ifftn.
 *   edbo_set_average, edbo_entropy_cal, edbo_fixsol
 *       golden values from the compiled EntropyFix.h (testEntropyFix inputs
 *       and seeded vectors).
 *   edbo_flogf, edbo_cal_entropy, edbo_bkw, edbo_grid_bkw
 *       closed forms (mass 1, zero momentum, energy 3; dK/dt consistency).
 *   edbo_run
 *       parity unpinned against the reference binary: BKW.h needs fftw3.h,
 *       which is absent, so the drivers cannot be built.  Pinned instead by
 *       each scheme's amplification polynomial on a linear operator and by
 *       convergence to the BKW exact solution.
 */
#include "edb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------- modes -- */

int edbo_nmodes(int N) { return 2 * ((N - 1) / 2) + 1; }

/* Wrap a frequency difference modulo N into the Galerkin range [-n, n].
 * For odd N every residue is a mode (QCalculator.h:204-206 moves the
 * out-of-range centred index back by N).  For even N the residue N/2 is the
 * Nyquist index, outside the Galerkin set: *valid = 0. */
int edbo_wrap(int N, int d, int* valid) {
  int n = (N - 1) / 2;
  int r = d % N;
  if (r < 0) r += N;
  *valid = 1;
  if (r <= n) return r;
  if (r >= N - n) return r - N;
  *valid = 0;
  return 0;
}

static int fftw_index(int N, int freq) { return freq >= 0 ? freq : freq + N; }

/* ------------------------------------------------------------- genB parts */

/* lambda = 2/(3+sqrt 2) * pi, QCalculator.h:163. */
double edbo_lambda(void) { return (2. / (3. + sqrt(2.))) * M_PI; }

/* F_alpha(xi, eta) = int_0^1 r^{2+alpha} sinc(xi r) sinc(eta r) dr in closed
 * form with the reference's special branches, QCalculator.h:76-106
 * (formulas ReadMe.md:51-68).  alpha outside {0,1} gives 0 like the
 * reference. */
double edbo_funF(int alpha, double xi, double eta) {
  double p = xi + eta, q = xi - eta;
  double res = 0.;
  if (alpha == 0) {
    if (eta < 1.0e-10) {
      if (xi < 1.0e-10) res = 1. / 3.;
      else res = (-xi * cos(xi) + sin(xi)) / (xi * xi * xi);
    } else {
      if (xi < 1.0e-10) res = (-eta * cos(eta) + sin(eta)) / (eta * eta * eta);
      else if (fabs(xi - eta) < 1.0e-10) res = (xi - cos(xi) * sin(xi)) / (2. * xi * xi * xi);
      else res = (p * sin(q) - q * sin(p)) / (2. * xi * eta * p * q);
    }
  } else if (alpha == 1) {
    if (eta < 1.0e-10) {
      if (xi < 1.0e-10) res = 1. / 4.;
      else res = (-2. - (-2. + xi * xi) * cos(xi) + 2. * xi * sin(xi)) / (xi * xi * xi * xi);
    } else {
      if (xi < 1.0e-10)
        res = (-2. - (-2. + eta * eta) * cos(eta) + 2. * eta * sin(eta)) / (eta * eta * eta * eta);
      else if (fabs(xi - eta) < 1.0e-10)
        res = -(-1. - 2. * xi * xi + cos(2. * xi) + 2. * xi * sin(2. * xi)) / (8. * xi * xi * xi * xi);
      else
        res = (q * sin(q) + cos(q)) / (2. * xi * eta * q * q) -
              (p * sin(p) + cos(p)) / (2. * xi * eta * p * p) - 2. / (p * p * q * q);
    }
  }
  return res;
}

/* Jackson filter sigma(beta) with n+1 = (N-1)/2 + 1, QCalculator.h:67-74. */
double edbo_filtering(int N, int beta) {
  int n = (N - 1) / 2;
  double tmp = (double)(n + 1);
  double PI = 4. * atan(1.);
  double b = fabs((double)beta);
  return ((tmp - b) * cos(PI * b / tmp) + sin(PI * b / tmp) / tan(PI / tmp)) / tmp;
}

/* rval = 4 pi (2 lambda / pi * L / N)^{3+alpha}, QCalculator.h:176 with the
 * hard-coded half-width 6.62132 generalised to L. */
double edbo_rval(int N, int alpha, double L) {
  double lambda = edbo_lambda();
  return (4. * M_PI) * pow(2. * lambda / M_PI * L / N, (double)(3 + alpha));
}

static double calxi(double lambda, int x1, int x2, int x3) {
  return sqrt((double)(x1 * x1 + x2 * x2 + x3 * x3)) * lambda; /* QCalculator.h:165-167 */
}

static double calfil(int N, int x1, int x2, int x3) {
  return edbo_filtering(N, x1) * edbo_filtering(N, x2) * edbo_filtering(N, x3); /* :168-170 */
}

/* beta(l, m) for frequency triples l, m (both Galerkin modes):
 * F(lambda|l+m|, lambda|l-m|) sigma_l sigma_m rval - F(lambda|2m|, 0) sigma_m^2 rval,
 * with the reference's operation order (QCalculator.h:181-192 for the
 * diagonal term Bmm, :221-224 for the entry). */
double edbo_beta(int N, int alpha, double L, int filter, int l1, int l2, int l3,
                 int m1, int m2, int m3) {
  double lambda = edbo_lambda();
  double rval = edbo_rval(N, alpha, L);
  double filv = 1., sigl = 1., sigm = 1.;
  if (filter) {
    filv = calfil(N, m1, m2, m3);
    sigl = calfil(N, l1, l2, l3);
    sigm = filv;
  }
  double bmm = edbo_funF(alpha, calxi(lambda, 2 * m1, 2 * m2, 2 * m3), 0.) * filv * filv * rval;
  double xi = calxi(lambda, l1 + m1, l2 + m2, l3 + m3);
  double eta = calxi(lambda, l1 - m1, l2 - m2, l3 - m3);
  return edbo_funF(alpha, xi, eta) * sigl * sigm * rval - bmm;
}

/* QCalculator::genB, QCalculator.h:158-229.  B has M^6 entries, row lc
 * (centred l) of M^3 entries indexed by centred k, holding
 * beta(l, [k-l]_N); 0 where [k-l]_N is not a Galerkin mode (even N). */
void edbo_genB(int N, int alpha, double L, int filter, double* B) {
  int n = (N - 1) / 2, M = edbo_nmodes(N);
  long NQ = (long)M * M * M;
  double lambda = edbo_lambda();
  double rval = edbo_rval(N, alpha, L);
  double* Bmm = (double*)malloc(sizeof(double) * NQ);
  double* sig = (double*)malloc(sizeof(double) * NQ);
  long ibm = 0;
  for (int k1 = -n; k1 < n + 1; ++k1)
    for (int k2 = -n; k2 < n + 1; ++k2)
      for (int k3 = -n; k3 < n + 1; ++k3) {
        double xi = calxi(lambda, 2 * k1, 2 * k2, 2 * k3);
        double fil = filter ? calfil(N, k1, k2, k3) : 1.;
        sig[ibm] = fil;
        Bmm[ibm++] = edbo_funF(alpha, xi, 0.) * fil * fil * rval;
      }
#pragma omp parallel for schedule(static)
  for (long lc = 0; lc < NQ; ++lc) {
    int l1 = (int)(lc / ((long)M * M)) - n, l2 = (int)((lc / M) % M) - n, l3 = (int)(lc % M) - n;
    double sigmal = sig[lc];
    double* row = B + lc * NQ;
    long ib = 0;
    for (int k1 = -n; k1 <= n; ++k1) {
      int v1, m1 = edbo_wrap(N, k1 - l1, &v1);
      for (int k2 = -n; k2 <= n; ++k2) {
        int v2, m2 = edbo_wrap(N, k2 - l2, &v2);
        for (int k3 = -n; k3 <= n; ++k3, ++ib) {
          int v3, m3 = edbo_wrap(N, k3 - l3, &v3);
          if (!(v1 && v2 && v3)) { row[ib] = 0.; continue; }
          long m = ((long)(m1 + n) * M + (m2 + n)) * M + (m3 + n);
          double xi = calxi(lambda, l1 + m1, l2 + m2, l3 + m3);
          double eta = calxi(lambda, l1 - m1, l2 - m2, l3 - m3);
          double sigmam = sig[m];
          row[ib] = edbo_funF(alpha, xi, eta) * sigmal * sigmam * rval - Bmm[m];
        }
      }
    }
  }
  free(Bmm);
  free(sig);
}

/* ----------------------------------------------------------------- calQ -- */

/* QCalculator::calQ (with transf/convertQ), QCalculator.h:108-156, 231-257:
 *   Q_k = sum_l (f_l * f_[k-l]) * B[l][k],  l ascending in centred order,
 * the complex product and the scaling by the real B done as std::complex
 * does them.  fin/qout are N^3 complex in FFTW order; qout at the Nyquist
 * index (even N) is 0.  The k range is split across threads; each k keeps the
 * reference's l order, so the result does not depend on nthreads. */
void edbo_calQ(int N, const double* B, const double* fin, double* qout, int nthreads) {
  int n = (N - 1) / 2, M = edbo_nmodes(N);
  long NQ = (long)M * M * M;
  /* f at centred index (transf duplicates f periodically; this gathers the
   * same values through explicit index maps). */
  double* fc = (double*)malloc(sizeof(double) * 2 * NQ);
  for (long c = 0; c < NQ; ++c) {
    int c1 = (int)(c / ((long)M * M)), c2 = (int)((c / M) % M), c3 = (int)(c % M);
    long j = ((long)fftw_index(N, c1 - n) * N + fftw_index(N, c2 - n)) * N + fftw_index(N, c3 - n);
    fc[2 * c] = fin[2 * j];
    fc[2 * c + 1] = fin[2 * j + 1];
  }
  /* wrap table for centred differences kc - lc in [-(M-1), M-1] */
  int* wt = (int*)malloc(sizeof(int) * (2 * M - 1));
  for (int d = -(M - 1); d <= M - 1; ++d) {
    int v, w = edbo_wrap(N, d, &v);
    wt[d + M - 1] = v ? w + n : -1;
  }
  double* Q = (double*)calloc(2 * NQ, sizeof(double));
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
  {
    int tid = 0, nth = 1;
#ifdef _OPENMP
    tid = omp_get_thread_num();
    nth = omp_get_num_threads();
#endif
    long k0 = NQ * tid / nth, k1e = NQ * (tid + 1) / nth;
    for (long lc = 0; lc < NQ; ++lc) {
      int l1 = (int)(lc / ((long)M * M)), l2 = (int)((lc / M) % M), l3 = (int)(lc % M);
      double ar = fc[2 * lc], ai = fc[2 * lc + 1];
      const double* brow = B + lc * NQ;
      for (long kc = k0; kc < k1e; ++kc) {
        int c1 = (int)(kc / ((long)M * M)), c2 = (int)((kc / M) % M), c3 = (int)(kc % M);
        int m1 = wt[c1 - l1 + M - 1], m2 = wt[c2 - l2 + M - 1], m3 = wt[c3 - l3 + M - 1];
        if (m1 < 0 || m2 < 0 || m3 < 0) continue;
        long m = ((long)m1 * M + m2) * M + m3;
        double cr = fc[2 * m], ci = fc[2 * m + 1];
        double b = brow[kc];
        double pr = ar * cr - ai * ci;
        double pi = ar * ci + ai * cr;
        Q[2 * kc] += pr * b;
        Q[2 * kc + 1] += pi * b;
      }
    }
  }
  /* convertQ: centred -> FFTW order, QCalculator.h:144-155 */
  memset(qout, 0, sizeof(double) * 2 * (long)N * N * N);
  for (long c = 0; c < NQ; ++c) {
    int c1 = (int)(c / ((long)M * M)), c2 = (int)((c / M) % M), c3 = (int)(c % M);
    long j = ((long)fftw_index(N, c1 - n) * N + fftw_index(N, c2 - n)) * N + fftw_index(N, c3 - n);
    qout[2 * j] = Q[2 * c];
    qout[2 * j + 1] = Q[2 * c + 1];
  }
  free(Q);
  free(wt);
  free(fc);
}

/* Direct evaluation in (l, m) order: Q_[l+m] += beta(l,m) f_l f_m over all
 * pairs of Galerkin modes whose wrapped sum is a mode.  This is the
 * "direct (l,m) layout" idea of oldTest/main.cpp:52-61 and the definition
 * the fast (l, k-l) layout must agree with. */
void edbo_calQ_direct(int N, int alpha, double L, int filter, const double* fin, double* qout) {
  int n = (N - 1) / 2;
  long N3 = (long)N * N * N;
  memset(qout, 0, sizeof(double) * 2 * N3);
  for (int l1 = -n; l1 <= n; ++l1)
    for (int l2 = -n; l2 <= n; ++l2)
      for (int l3 = -n; l3 <= n; ++l3) {
        long jl = ((long)fftw_index(N, l1) * N + fftw_index(N, l2)) * N + fftw_index(N, l3);
        double ar = fin[2 * jl], ai = fin[2 * jl + 1];
        for (int m1 = -n; m1 <= n; ++m1)
          for (int m2 = -n; m2 <= n; ++m2)
            for (int m3 = -n; m3 <= n; ++m3) {
              int v1, v2, v3;
              int k1 = edbo_wrap(N, l1 + m1, &v1), k2 = edbo_wrap(N, l2 + m2, &v2),
                  k3 = edbo_wrap(N, l3 + m3, &v3);
              if (!(v1 && v2 && v3)) continue;
              long jm = ((long)fftw_index(N, m1) * N + fftw_index(N, m2)) * N + fftw_index(N, m3);
              long jk = ((long)fftw_index(N, k1) * N + fftw_index(N, k2)) * N + fftw_index(N, k3);
              double b = edbo_beta(N, alpha, L, filter, l1, l2, l3, m1, m2, m3);
              double cr = fin[2 * jm], ci = fin[2 * jm + 1];
              qout[2 * jk] += (ar * cr - ai * ci) * b;
              qout[2 * jk + 1] += (ar * ci + ai * cr) * b;
            }
      }
}

/* Q at selected modes without the M^6 array: for each centred mode index
 * kc[i], Q_k = sum_l (f_l * f_[k-l]) * beta(l, [k-l]) in edbo_calQ's l order
 * and arithmetic, with each beta formed by edbo_beta (the per-entry formula
 * of QCalculator.h:199-228).  absval != 0 accumulates |beta| |f_l| |f_[k-l]|
 * into the real part instead (the scale of the summed terms).  fin: N^3
 * complex in FFTW order; out: nk complex.  Cost M^3 beta evaluations per
 * mode, so it serves grids whose M^6 array would not fit. */
void edbo_calQ_modes(int N, int alpha, double L, int filter, const double* fin, int nk, const long* kc,
                     int absval, double* out) {
  int n = (N - 1) / 2, M = edbo_nmodes(N);
  long NQ = (long)M * M * M;
  double* fc = (double*)malloc(sizeof(double) * 2 * NQ);
  for (long c = 0; c < NQ; ++c) {
    int c1 = (int)(c / ((long)M * M)), c2 = (int)((c / M) % M), c3 = (int)(c % M);
    long j = ((long)fftw_index(N, c1 - n) * N + fftw_index(N, c2 - n)) * N + fftw_index(N, c3 - n);
    fc[2 * c] = fin[2 * j];
    fc[2 * c + 1] = fin[2 * j + 1];
  }
  int* wt = (int*)malloc(sizeof(int) * (2 * M - 1));
  for (int d = -(M - 1); d <= M - 1; ++d) {
    int v, w = edbo_wrap(N, d, &v);
    wt[d + M - 1] = v ? w + n : -1;
  }
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nk; ++i) {
    long k = kc[i];
    int c1 = (int)(k / ((long)M * M)), c2 = (int)((k / M) % M), c3 = (int)(k % M);
    double qr = 0., qi = 0.;
    for (long lc = 0; lc < NQ; ++lc) {
      int l1 = (int)(lc / ((long)M * M)), l2 = (int)((lc / M) % M), l3 = (int)(lc % M);
      int m1 = wt[c1 - l1 + M - 1], m2 = wt[c2 - l2 + M - 1], m3 = wt[c3 - l3 + M - 1];
      if (m1 < 0 || m2 < 0 || m3 < 0) continue;
      long m = ((long)m1 * M + m2) * M + m3;
      double ar = fc[2 * lc], ai = fc[2 * lc + 1];
      double cr = fc[2 * m], ci = fc[2 * m + 1];
      double b = edbo_beta(N, alpha, L, filter, l1 - n, l2 - n, l3 - n, m1 - n, m2 - n, m3 - n);
      if (absval) {
        qr += fabs(b) * sqrt(ar * ar + ai * ai) * sqrt(cr * cr + ci * ci);
      } else {
        double pr = ar * cr - ai * ci;
        double pi = ar * ci + ai * cr;
        qr += pr * b;
        qi += pi * b;
      }
    }
    out[2 * i] = qr;
    out[2 * i + 1] = qi;
  }
  free(wt);
  free(fc);
}

/* ------------------------------------------------------------------ DFT -- */

/* Separable textbook DFT, out_k = sum_j in_j exp(sign 2 pi i jk/N) along
 * each axis, unnormalised: the FFTW_FORWARD (sign -1) / FFTW_BACKWARD (+1)
 * conventions that fft3d.h:25-30,32-37 select. */
void edbo_dft3(int N, int sign, const double* in, double* out) {
  long N3 = (long)N * N * N;
  double* cw = (double*)malloc(sizeof(double) * N);
  double* sw = (double*)malloc(sizeof(double) * N);
  for (int r = 0; r < N; ++r) {
    double a = 2. * M_PI * (double)r / (double)N;
    cw[r] = cos(a);
    sw[r] = sign * sin(a);
  }
  double* a = (double*)malloc(sizeof(double) * 2 * N3);
  double* b = (double*)malloc(sizeof(double) * 2 * N3);
  memcpy(a, in, sizeof(double) * 2 * N3);
  long stride[3] = {(long)N * N, N, 1};
  for (int ax = 0; ax < 3; ++ax) {
    long s = stride[ax];
    for (long base = 0; base < N3; ++base) {
      if ((base / s) % N != 0) continue; /* base has index 0 along ax */
      for (int k = 0; k < N; ++k) {
        double sr = 0., si = 0.;
        for (int j = 0; j < N; ++j) {
          int r = (int)(((long)j * k) % N);
          double xr = a[2 * (base + j * s)], xi = a[2 * (base + j * s) + 1];
          sr += xr * cw[r] - xi * sw[r];
          si += xr * sw[r] + xi * cw[r];
        }
        b[2 * (base + k * s)] = sr;
        b[2 * (base + k * s) + 1] = si;
      }
    }
    double* t = a; a = b; b = t;
  }
  memcpy(out, a, sizeof(double) * 2 * N3);
  free(a); free(b); free(cw); free(sw);
}

/* fft3d::fft, fft3d.h:25-30 (forward, unnormalised). */
void edbo_fft(int N, const double* re, double* fre) { edbo_dft3(N, -1, re, fre); }

/* fft3d::ifft, fft3d.h:32-48: backward transform then each component times
 * 1./N^3. */
void edbo_ifft(int N, double* re, const double* fre) {
  long N3 = (long)N * N * N;
  edbo_dft3(N, +1, fre, re);
  double s = 1. / (double)N3;
  for (long k = 0; k < 2 * N3; ++k) re[k] *= s;
}

/* ------------------------------------------ conservative correction -- */

/* Conservative moment correction of a collision operator's spectrum.  No
 * counterpart in the reference (its Q conserves mass only, through
 * Q_0 == 0).  The discrete moments m_i(g) = h^3 sum_x phi_i(x) g(x) over the
 * grid, phi in {1, v1, v2, v3, |v|^2}, are linear functionals of the
 * spectrum: with P_i = DFT(phi_i) (forward, unnormalised),
 *   m_i(g) = h^3/N^3 sum_k ghat_k conj(P_i,k).
 * The corrected spectrum is the least-squares nearest one (discrete L2 norm
 * over the Galerkin modes G) with all five moments zero:
 *   Qhat' = Qhat - sum_j mu_j P_j (on G),   A mu = m,   A_ij = sum_{k in G} P_j,k conj(P_i,k).
 * For odd N (G = every mode) this is the constrained L2 projection of the
 * grid values, Q' = Q - C^T (C C^T)^-1 C Q.  Restated here with P computed by
 * the O(N^4) DFT of the five grid functions; Qhat (FFTW order) in place. */
static double* mc_P = NULL; /* P of the last (N, L), kept between calls */
static int mc_N = 0;
static double mc_L = 0.;

void edbo_moment_correct(int N, double L, double* Qhat) {
  long N3 = (long)N * N * N;
  double h = 2. * L / (double)N;
  double* g = (double*)malloc(sizeof(double) * 2 * N3);
  if (mc_P && (mc_N != N || mc_L != L)) {
    free(mc_P);
    mc_P = NULL;
  }
  int fresh = mc_P == NULL;
  if (fresh) mc_P = (double*)malloc(sizeof(double) * 2 * N3 * 5);
  double* P = mc_P;
  mc_N = N;
  mc_L = L;
  for (int i = 0; i < 5 && fresh; ++i) {
    for (long x = 0; x < N3; ++x) {
      int j1 = (int)(x / ((long)N * N)), j2 = (int)((x / N) % N), j3 = (int)(x % N);
      double v1 = -L + j1 * h, v2 = -L + j2 * h, v3 = -L + j3 * h;
      double phi = i == 0 ? 1. : i == 1 ? v1 : i == 2 ? v2 : i == 3 ? v3 : v1 * v1 + v2 * v2 + v3 * v3;
      g[2 * x] = phi;
      g[2 * x + 1] = 0.;
    }
    edbo_fft(N, g, P + 2 * N3 * i);
    if (N % 2 == 0) /* Nyquist planes are outside G */
      for (long x = 0; x < N3; ++x) {
        int j1 = (int)(x / ((long)N * N)), j2 = (int)((x / N) % N), j3 = (int)(x % N);
        if (2 * j1 == N || 2 * j2 == N || 2 * j3 == N) P[2 * (N3 * i + x)] = P[2 * (N3 * i + x) + 1] = 0.;
      }
  }
  /* A (real: G is symmetric under k -> -k and P_i,-k = conj(P_i,k)), m complex */
  double A[5][5], mr[5], mi[5];
  for (int i = 0; i < 5; ++i) {
    mr[i] = mi[i] = 0.;
    for (long x = 0; x < N3; ++x) {
      double pr = P[2 * (N3 * i + x)], pi = -P[2 * (N3 * i + x) + 1]; /* conj(P_i) */
      double qr = Qhat[2 * x], qi = Qhat[2 * x + 1];
      mr[i] += qr * pr - qi * pi;
      mi[i] += qr * pi + qi * pr;
    }
    for (int j = 0; j < 5; ++j) {
      double a = 0.;
      for (long x = 0; x < N3; ++x)
        a += P[2 * (N3 * j + x)] * P[2 * (N3 * i + x)] + P[2 * (N3 * j + x) + 1] * P[2 * (N3 * i + x) + 1];
      A[i][j] = a;
    }
  }
  /* Gaussian elimination with partial pivoting on the diagonally scaled
   * system (the functionals differ by powers of L). */
  double d[5], M[5][7];
  for (int i = 0; i < 5; ++i) d[i] = 1. / sqrt(A[i][i]);
  for (int i = 0; i < 5; ++i) {
    for (int j = 0; j < 5; ++j) M[i][j] = d[i] * A[i][j] * d[j];
    M[i][5] = d[i] * mr[i];
    M[i][6] = d[i] * mi[i];
  }
  for (int c = 0; c < 5; ++c) {
    int piv = c;
    for (int r = c + 1; r < 5; ++r)
      if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
    for (int j = 0; j < 7; ++j) {
      double t = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = t;
    }
    for (int r = 0; r < 5; ++r) {
      if (r == c) continue;
      double f = M[r][c] / M[c][c];
      for (int j = c; j < 7; ++j) M[r][j] -= f * M[c][j];
    }
  }
  double mur[5], mui[5];
  for (int i = 0; i < 5; ++i) {
    mur[i] = d[i] * M[i][5] / M[i][i];
    mui[i] = d[i] * M[i][6] / M[i][i];
  }
  for (long x = 0; x < N3; ++x)
    for (int j = 0; j < 5; ++j) {
      double pr = P[2 * (N3 * j + x)], pi = P[2 * (N3 * j + x) + 1];
      Qhat[2 * x] -= mur[j] * pr - mui[j] * pi;
      Qhat[2 * x + 1] -= mur[j] * pi + mui[j] * pr;
    }
  free(g);
}

/* ---------------------------------------------------------- EntropyFix -- */

/* EntropyFix(Nt, eps), EntropyFix.h:29-32; Averagef = 1.0 until set, :20. */
void edbo_efix_init(edbo_efix* e, long Nt, double eps) {
  e->Nt = Nt;
  e->eps = eps;
  e->average = 1.0;
  e->average_flag = 0;
}

/* setAverage, EntropyFix.h:65-78: mean of f, summed in index order. */
void edbo_set_average(edbo_efix* e, long Nf, const double* f) {
  e->average_flag = 1;
  double a = 0.;
  for (long i = 0; i < Nf; ++i) a += f[i];
  e->average = a / (double)Nf;
}

/* EntropyCal, EntropyFix.h:124-161: (sum_{f > eps} f log f)/Nf - Averagef,
 * the average cached at the first call. */
double edbo_entropy_cal(edbo_efix* e, long Nf, const double* f) {
  double res = 0.0;
  if (!e->average_flag) edbo_set_average(e, Nf, f);
  for (long i = 0; i < Nf; i++)
    if (f[i] > e->eps) res += f[i] * log(f[i]);
  return res / (double)Nf - e->average;
}

/* FixSol(Nf, fin, Aim), EntropyFix.h:211-253: bisection on beta in [0,1]
 * for f_beta = (1-beta) f + beta*Averagef until width <= eps; early exit
 * when Aim-eps < H(f_beta) < Aim, else f <- f_{betaR}. */
void edbo_fixsol(edbo_efix* e, long Nf, double* fin, double aim) {
  double* fMid = (double*)malloc(sizeof(double) * Nf);
  double betaL = 0.0, betaR = 1.0, betaMid, entMid, betaFbar;
  double BiEps = e->eps;
  while ((betaR - betaL) > BiEps) {
    betaMid = 0.5 * (betaL + betaR);
    if (!e->average_flag) edbo_set_average(e, Nf, fin);
    betaFbar = betaMid * e->average;
    for (long i = 0; i < Nf; i++) fMid[i] = betaFbar;
    for (long i = 0; i < Nf; ++i) fMid[i] += (1. - betaMid) * fin[i];
    entMid = edbo_entropy_cal(e, Nf, fMid);
    if ((entMid < aim) && ((aim - entMid) < e->eps)) {
      for (long i = 0; i < Nf; ++i) fin[i] = fMid[i];
      free(fMid);
      return;
    }
    if (entMid < aim) betaR = betaMid;
    else betaL = betaMid;
  }
  betaFbar = betaR * e->average;
  for (long i = 0; i < Nf; ++i) fin[i] *= (1.0 - betaR);
  for (long i = 0; i < Nf; i++) fin[i] += betaFbar;
  free(fMid);
}

/* ------------------------------------------------------------- BKW3D ---- */

/* BKW3D::flogf, BKW.h:153-161. */
double edbo_flogf(double f) {
  if (fabs(f) < 1.0e-6) return 0.;
  return f * log(f) - f;
}

/* BKW3D::calEntropy, BKW.h:140-152: (2R)^3/N^3 sum flogf(Re F). */
double edbo_cal_entropy(int N, double R, const double* Fc) {
  long N3 = (long)N * N * N;
  double entropy = 0;
  for (long i = 0; i < N3; ++i) entropy += edbo_flogf(Fc[2 * i]);
  entropy *= pow(2 * R, 3.) / (double)N3;
  return entropy;
}

/* BKW exact solution with K = 1 - exp(-t/6), the live copy at
 * testfft.cpp:11-16 (dead branch BKW.h:113-117) written in K:
 * (2 pi K)^{-3/2} exp(-v^2/2K) ((5K-3)/2K + (1-K)/(2K^2) v^2). */
double edbo_bkw(double K, double x, double y, double z) {
  double v2 = x * x + y * y + z * z;
  return 1. / pow(2. * M_PI * K, 1.5) * exp(-v2 / (2. * K)) *
         ((5. * K - 3.) / (2. * K) + (1. - K) / (2. * K * K) * v2);
}

/* BKW3D::setInitialValue, BKW.h:121-138: x_i = -R + i*2R/N, real data in
 * complex storage, with the BKW profile at K. */
void edbo_grid_bkw(int N, double R, double K, double* Fc) {
  double h = 2 * R / (double)N;
  long l = 0;
  for (int i = 0; i < N; ++i) {
    double x = -R + i * h;
    for (int j = 0; j < N; ++j) {
      double y = -R + j * h;
      for (int k = 0; k < N; ++k) {
        double z = -R + k * h;
        Fc[2 * l] = edbo_bkw(K, x, y, z);
        Fc[2 * l + 1] = 0.;
        l++;
      }
    }
  }
}

/* ----------------------------------------------------------- drivers ---- */

/* The explicit drivers of BKW3D (BKW.h:191-633) restated on the spectral
 * state: stage combinations in the reference's operation order, then ifft,
 * calEntropy and the entropy fix.  The reference's FixSol result only lands
 * in the real-space F (diagnostic-only, BKW.h:225-235, 497-510); fix_mode 2
 * also transforms the fixed F back into Fhat as the disabled branches at
 * BKW.h:358-376 do.  Fix eligibility follows each driver: FEuler from its
 * second step (BKW.h:227-229), CFE/SSP3/RK4 from the second step of the loop
 * (`if(step)`), RK2/DUFORT never in mode 1 (`if(0)`). */
typedef struct {
  int N;
  long N3;
  double R, vol;
  double *F, *Ftmp;
  edbo_efix efix;
  int fix_mode;
  double* Fhat;
  int rec, max_steps;
  double *t_out, *H_out, *Hfix_out;
} post_ctx;

static double post_step(post_ctx* p, double t, int eligible, double entropy0) {
  long N3 = p->N3;
  edbo_ifft(p->N, p->F, p->Fhat);
  double entropy = edbo_cal_entropy(p->N, p->R, p->F);
  double h_before = entropy;
  if (eligible && p->fix_mode != 0 && entropy - entropy0 > 1.0e-10) {
    for (long i = 0; i < N3; ++i) p->Ftmp[i] = p->F[2 * i];
    edbo_fixsol(&p->efix, N3, p->Ftmp, entropy0 / p->vol);
    for (long i = 0; i < N3; ++i) { p->F[2 * i] = p->Ftmp[i]; p->F[2 * i + 1] = 0.; }
    entropy = edbo_cal_entropy(p->N, p->R, p->F);
    if (p->fix_mode == 2) edbo_fft(p->N, p->F, p->Fhat);
  }
  if (p->rec < p->max_steps) {
    p->t_out[p->rec] = t;
    p->H_out[p->rec] = h_before;
    p->Hfix_out[p->rec] = entropy;
  }
  p->rec++;
  return entropy;
}

/* calQ of the driver, followed by the conservative correction when asked. */
static void drv_calQ(int N, const double* B, double R, int conserve, const double* F, double* Q,
                     int nthreads) {
  edbo_calQ(N, B, F, Q, nthreads);
  if (conserve) edbo_moment_correct(N, R, Q);
}

int edbo_run(int N, const double* B, double R, int scheme, double T0, double TEnd, double dt,
             double beta_cfe, int fix_mode, double* Fhat, int max_steps, double* t_out,
             double* H_out, double* Hfix_out, int nthreads) {
  return edbo_run_ex(N, B, R, scheme, T0, TEnd, dt, beta_cfe, fix_mode, 0, Fhat, max_steps, t_out,
                     H_out, Hfix_out, nthreads);
}

int edbo_run_ex(int N, const double* B, double R, int scheme, double T0, double TEnd, double dt,
                double beta_cfe, int fix_mode, int conserve, double* Fhat, int max_steps,
                double* t_out, double* H_out, double* Hfix_out, int nthreads) {
  long N3 = (long)N * N * N, n2 = 2 * N3;
  size_t cb = sizeof(double) * 2 * N3;
  double *Q = malloc(cb), *Q1 = malloc(cb), *Q2 = malloc(cb), *Q3 = malloc(cb);
  double *F1 = malloc(cb), *F2 = malloc(cb), *F3 = malloc(cb), *Fold = malloc(cb);
  post_ctx p;
  p.N = N; p.N3 = N3; p.R = R; p.vol = pow(2 * R, 3);
  p.F = malloc(cb); p.Ftmp = malloc(sizeof(double) * N3);
  edbo_efix_init(&p.efix, N3, 1.0e-10);
  p.fix_mode = fix_mode; p.Fhat = Fhat; p.rec = 0; p.max_steps = max_steps;
  p.t_out = t_out; p.H_out = H_out; p.Hfix_out = Hfix_out;
  int fixable = !(fix_mode == 1 && (scheme == 1 || scheme == 5));
  double t = T0, entropy0 = 1.;
  int step = 0;

  if (scheme == 0) { /* FEuler, BKW.h:214-248 */
    edbo_ifft(N, p.F, Fhat);
    entropy0 = edbo_cal_entropy(N, R, p.F);
    step = 1;
    do {
      t += dt;
      drv_calQ(N, B, R, conserve, Fhat, Q, nthreads);
      for (long l = 0; l < n2; ++l) Fhat[l] = Fhat[l] + dt * Q[l];
      entropy0 = post_step(&p, t, step > 1, entropy0);
      step++;
    } while (t < TEnd);
  } else if (scheme == 4) { /* CFE, BKW.h:277-311 */
    double s = (1. - exp(-beta_cfe * dt)) / beta_cfe;
    do {
      t += dt;
      drv_calQ(N, B, R, conserve, Fhat, Q, nthreads);
      for (long l = 0; l < n2; ++l) Fhat[l] = Fhat[l] + s * Q[l];
      entropy0 = post_step(&p, t, step > 0, entropy0);
      step++;
    } while (t < TEnd);
  } else if (scheme == 1 || scheme == 5) { /* RK2 BKW.h:337-382; DUFORT start-up :546-587 */
    if (scheme == 5) memcpy(Fold, Fhat, cb); /* BKW.h:539 */
    double tstep = scheme == 5 ? dt / 8 : dt;
    do {
      t += tstep;
      drv_calQ(N, B, R, conserve, Fhat, Q, nthreads);
      for (long l = 0; l < n2; ++l) F1[l] = Fhat[l] + 0.5 * dt * Q[l];
      drv_calQ(N, B, R, conserve, F1, Q, nthreads);
      for (long l = 0; l < n2; ++l) Fhat[l] = Fhat[l] + dt * Q[l];
      entropy0 = post_step(&p, t, fixable && step > 0, entropy0);
      step++;
    } while (scheme == 5 ? step < 8 : t < TEnd);
    if (scheme == 5) { /* DuFort-Frankel, BKW.h:592-631 */
      step = 0;
      do {
        t += dt;
        drv_calQ(N, B, R, conserve, Fhat, Q, nthreads);
        memcpy(F1, Fhat, cb);
        for (long l = 0; l < n2; ++l)
          Fhat[l] = 1. / (1. + beta_cfe * dt) * (F1[l] + dt * Q[l] + beta_cfe * dt * (F1[l] - Fold[l]));
        entropy0 = post_step(&p, t, fixable && step > 0, entropy0);
        step++;
        memcpy(Fold, F1, cb);
      } while (t < TEnd);
    }
  } else if (scheme == 2) { /* SSP3, BKW.h:413-446 */
    do {
      t += dt;
      drv_calQ(N, B, R, conserve, Fhat, Q, nthreads);
      for (long l = 0; l < n2; ++l) F1[l] = Fhat[l] + dt * Q[l];
      drv_calQ(N, B, R, conserve, F1, Q, nthreads);
      for (long l = 0; l < n2; ++l) F2[l] = 0.75 * Fhat[l] + 0.25 * F1[l] + 0.25 * dt * Q[l];
      drv_calQ(N, B, R, conserve, F2, Q, nthreads);
      for (long l = 0; l < n2; ++l) Fhat[l] = Fhat[l] / 3. + 2. * F2[l] / 3. + 2. / 3. * dt * Q[l];
      entropy0 = post_step(&p, t, step > 0, entropy0);
      step++;
    } while (t < TEnd);
  } else { /* RK4, BKW.h:478-516 */
    do {
      t += dt;
      drv_calQ(N, B, R, conserve, Fhat, Q, nthreads);
      for (long l = 0; l < n2; ++l) F1[l] = Fhat[l] + 0.5 * dt * Q[l];
      drv_calQ(N, B, R, conserve, F1, Q1, nthreads);
      for (long l = 0; l < n2; ++l) F2[l] = Fhat[l] + 0.5 * dt * Q1[l];
      drv_calQ(N, B, R, conserve, F2, Q2, nthreads);
      for (long l = 0; l < n2; ++l) F3[l] = Fhat[l] + dt * Q2[l];
      drv_calQ(N, B, R, conserve, F3, Q3, nthreads);
      for (long l = 0; l < n2; ++l)
        Fhat[l] = Fhat[l] + dt / 6. * (Q[l] + 2. * Q1[l] + 2. * Q2[l] + Q3[l]);
      entropy0 = post_step(&p, t, step > 0, entropy0);
      step++;
    } while (t < TEnd);
  }
  free(p.F); free(p.Ftmp);
  free(Q); free(Q1); free(Q2); free(Q3); free(F1); free(F2); free(F3); free(Fold);
  return p.rec;
}
