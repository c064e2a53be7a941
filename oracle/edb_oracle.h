/*
 * edb_oracle.h -- CPU oracle for the CalQ collision-operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (YK2015/EntropyDecreaseBoltzmann, the CalQ headers) that the HIP
 * library's results are checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time the CPU
 * baseline.  The product path never calls into it.
 *
 * Conventions (shared with the HIP library):
 *   N      FFT grid points per dimension, velocity domain [-L, L]^3,
 *          x_i = -L + i*2L/N (CalQ/BKW.h:121-138).
 *   n      (N-1)/2 (integer division), QCalculator.h:44.
 *   M      2n+1 Fourier-Galerkin modes per dimension, frequencies -n..n.
 *          For odd N, M == N and everything below reduces to the reference
 *          bit for bit.  For even N the reference is undefined (its calQ walks
 *          N-1 k values against rows laid out with N); here the Galerkin set
 *          is the same 2n+1 = N-1 modes, differences wrap modulo N, and a
 *          term whose wrapped partner falls on the Nyquist index is absent.
 *   FFTW order     index j <-> freq j (j <= n) or j-N (j > n); for even N the
 *                  Nyquist index N/2 is outside the Galerkin set.
 *   centred order  index c <-> freq c-n, c in [0, M).
 *   complex arrays interleaved (re, im) doubles, like fftw_complex.
 *   B layout       B[lc*M^3 + kc] = beta(l, [k-l]_N), l and k centred,
 *                  row-major (QCalculator.h:199-228); 0 when [k-l]_N is not a
 *                  Galerkin mode.
 *
 * Pinning (tests/test_oracle.py): golden vectors written by the compiled
 * reference headers (QCalculator.h, EntropyFix.h) for odd N, brute-force
 * direct sums, numpy.fft, closed forms of F_alpha, invariants (Q_0 == 0,
 * Hermitian symmetry), and the BKW exact solution.
 */
#ifndef EDB_ORACLE_H_
#define EDB_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

int    edbo_nmodes(int N);
double edbo_lambda(void);
double edbo_funF(int alpha, double xi, double eta);
double edbo_filtering(int N, int beta);
double edbo_rval(int N, int alpha, double L);
int    edbo_wrap(int N, int d, int* valid);
double edbo_beta(int N, int alpha, double L, int filter,
                 int l1, int l2, int l3, int m1, int m2, int m3);
void   edbo_genB(int N, int alpha, double L, int filter, double* B);
void   edbo_calQ(int N, const double* B, const double* fin, double* qout,
                 int nthreads);
void   edbo_calQ_direct(int N, int alpha, double L, int filter,
                        const double* fin, double* qout);
void   edbo_calQ_modes(int N, int alpha, double L, int filter, const double* fin, int nk,
                       const long* kc, int absval, double* out);

void   edbo_dft3(int N, int sign, const double* in, double* out);
void   edbo_fft(int N, const double* re, double* fre);
void   edbo_ifft(int N, double* re, const double* fre);

typedef struct {
  long   Nt;
  double eps;
  double average;
  int    average_flag;
} edbo_efix;

void   edbo_efix_init(edbo_efix* e, long Nt, double eps);
void   edbo_set_average(edbo_efix* e, long Nf, const double* f);
double edbo_entropy_cal(edbo_efix* e, long Nf, const double* f);
void   edbo_fixsol(edbo_efix* e, long Nf, double* f, double aim);

double edbo_flogf(double f);
double edbo_cal_entropy(int N, double R, const double* Fc);
double edbo_bkw(double K, double x, double y, double z);
void   edbo_grid_bkw(int N, double R, double K, double* Fc);

/* Explicit schemes of BKW3D (BKW.h:191-633) on the spectral state.
 * scheme: 0 FEuler, 1 RK2, 2 SSP3, 3 RK4, 4 CFE, 5 DUFORT.
 * fix_mode: 0 none, 1 reference diagnostic-only fix (real-space F fixed and
 * reported, spectral state untouched), 2 applied fix (fixed F transformed
 * back into Fhat).
 * Fhat (N^3 complex) is advanced in place.  Per step k (0-based) it records
 * t[k], the entropy after the step H[k] and the entropy after the fix Hfix[k]
 * (== H[k] when no fix ran) into the caller's arrays of length max_steps.
 * Returns the number of steps taken (the reference's do-while count). */
int    edbo_run(int N, const double* B, double R, int scheme, double T0,
                double TEnd, double dt, double beta_cfe, int fix_mode,
                double* Fhat, int max_steps, double* t_out,
                double* H_out, double* Hfix_out, int nthreads);
/* edbo_run with every Q passed through edbo_moment_correct when conserve. */
int    edbo_run_ex(int N, const double* B, double R, int scheme, double T0,
                   double TEnd, double dt, double beta_cfe, int fix_mode,
                   int conserve, double* Fhat, int max_steps, double* t_out,
                   double* H_out, double* Hfix_out, int nthreads);

/* Conservative moment correction (no counterpart in the reference): the
 * least-squares nearest spectrum on the Galerkin modes whose discrete mass,
 * momentum and energy moments vanish.  Qhat N^3 complex, FFTW order, in
 * place; L the velocity half-width. */
void   edbo_moment_correct(int N, double L, double* Qhat);

#ifdef __cplusplus
}
#endif
#endif
