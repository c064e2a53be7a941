/* Drives every entry point of the CPU oracle on small odd and even grids,
 * with light consistency checks, so that a build with
 * -fsanitize=address,undefined exercises all of its memory accesses and
 * integer arithmetic.  Exit status 0 when every check holds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "edb_oracle.h"

static int fails = 0;

static void expect(int ok, const char* what, int N) {
  if (!ok) {
    fprintf(stderr, "FAIL %s (N=%d)\n", what, N);
    fails++;
  }
}

static double maxabs(const double* a, long n) {
  double m = 0.;
  for (long i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
  return m;
}

int main(void) {
  const double L = 8.0;
  for (int N = 3; N <= 8; ++N) {
    const int M = edbo_nmodes(N), n = (N - 1) / 2;
    const long N3 = (long)N * N * N, M3 = (long)M * M * M;
    double* B = malloc(sizeof(double) * M3 * M3);
    double* F = malloc(sizeof(double) * 2 * N3);
    double* Fh = malloc(sizeof(double) * 2 * N3);
    double* G = malloc(sizeof(double) * 2 * N3);
    double* Q = malloc(sizeof(double) * 2 * N3);
    double* Qd = malloc(sizeof(double) * 2 * N3);
    edbo_grid_bkw(N, L, 0.6, F);
    edbo_fft(N, F, Fh);
    edbo_ifft(N, G, Fh);
    double err = 0.;
    for (long i = 0; i < 2 * N3; ++i) err = fmax(err, fabs(G[i] - F[i]));
    expect(err <= 1e-12 * maxabs(F, 2 * N3), "fft round trip", N);
    edbo_dft3(N, -1, F, G);
    expect(G[0] == Fh[0] && G[1] == Fh[1], "dft3 equals fft", N);
    for (int alpha = 0; alpha <= 1; ++alpha)
      for (int filter = 0; filter <= 1; ++filter) {
        edbo_genB(N, alpha, L, filter, B);
        edbo_calQ(N, B, Fh, Q, 2);
        edbo_calQ_direct(N, alpha, L, filter, Fh, Qd);
        double d = 0.;
        for (long i = 0; i < 2 * N3; ++i) d = fmax(d, fabs(Q[i] - Qd[i]));
        expect(d <= 1e-12 * fmax(maxabs(Q, 2 * N3), 1e-300), "calQ against the direct sum", N);
        expect(Q[0] == 0. && Q[1] == 0., "Q_0 == 0", N);
        long kc[3] = {0, M3 / 2, M3 - 1};
        double out[6], outa[6];
        edbo_calQ_modes(N, alpha, L, filter, Fh, 3, kc, 0, out);
        edbo_calQ_modes(N, alpha, L, filter, Fh, 3, kc, 1, outa);
        for (int i = 0; i < 3; ++i) expect(hypot(out[2 * i], out[2 * i + 1]) <= outa[2 * i] * (1 + 1e-12) + 1e-300,
                                           "|Q_k| bounded by its term sum", N);
        /* the centre mode (k = 0) of calQ_modes is Q_0 */
        expect(out[2] == 0. && out[3] == 0., "calQ_modes at k = 0", N);
        double beta = edbo_beta(N, alpha, L, filter, 1 % (n + 1), 0, 0, 0, 1 % (n + 1), 0);
        expect(isfinite(beta), "beta entry", N);
        for (long i = 0; i < 2 * N3; ++i) G[i] = Q[i];
        edbo_moment_correct(N, L, G);
        expect(isfinite(maxabs(G, 2 * N3)), "moment correction", N);
      }
    int valid = 0;
    expect(edbo_wrap(N, N - 1, &valid) == -1 || !valid, "wrap", N);
    expect(edbo_filtering(N, 0) == 1., "filtering(0)", N);
    expect(isfinite(edbo_funF(0, 1.5, 0.5)) && isfinite(edbo_funF(1, 0., 0.)), "funF", N);
    expect(edbo_rval(N, 0, L) > 0. && edbo_lambda() > 0., "rval, lambda", N);
    expect(isfinite(edbo_cal_entropy(N, L, F)), "calEntropy", N);
    expect(edbo_flogf(0.) == 0., "flogf", N);
    /* entropy fix on the real part of F */
    double* f = malloc(sizeof(double) * N3);
    for (long i = 0; i < N3; ++i) f[i] = F[2 * i] + 1e-3;
    edbo_efix e;
    edbo_efix_init(&e, N3, 1e-10);
    edbo_set_average(&e, N3, f);
    const double H = edbo_entropy_cal(&e, N3, f);
    double mass0 = 0., mass1 = 0.;
    for (long i = 0; i < N3; ++i) mass0 += f[i];
    edbo_fixsol(&e, N3, f, H - 1e-3 * fabs(H));
    for (long i = 0; i < N3; ++i) mass1 += f[i];
    expect(fabs(mass1 - mass0) <= 1e-12 * fabs(mass0), "FixSol keeps the mass", N);
    free(f);
    /* the six drivers, each fix mode, with and without the correction */
    edbo_genB(N, 0, L, 0, B);
    for (int scheme = 0; scheme <= 5; ++scheme)
      for (int fix = 0; fix <= 2; ++fix)
        for (int conserve = 0; conserve <= 1; ++conserve) {
          double t[64], Hs[64], Hf[64];
          for (long i = 0; i < 2 * N3; ++i) G[i] = Fh[i];
          const int steps = edbo_run_ex(N, B, L, scheme, 0., 0.03, 0.01, 1.0, fix, conserve, G, 64, t, Hs, Hf, 2);
          expect(steps >= 3 && steps <= 64, "driver step count", N);
          expect(isfinite(maxabs(G, 2 * N3)), "driver state", N);
        }
    {
      double t[64], Hs[64], Hf[64];
      for (long i = 0; i < 2 * N3; ++i) G[i] = Fh[i];
      expect(edbo_run(N, B, L, 3, 0., 0.02, 0.01, 0., 1, G, 64, t, Hs, Hf, 1) >= 2, "edbo_run", N);
    }
    free(B);
    free(F);
    free(Fh);
    free(G);
    free(Q);
    free(Qd);
  }
  if (fails) fprintf(stderr, "%d checks failed\n", fails);
  else printf("oracle sanitizer driver: all checks passed\n");
  return fails ? 1 : 0;
}
