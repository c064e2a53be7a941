// genb.hip -- the collision kernel beta on the GPU, in the three layouts the
// Q kernels stream.
//
// beta(l, m) = F_alpha(lambda|l+m|, lambda|l-m|) sigma_l sigma_m rval
//            - F_alpha(lambda|2m|, 0) sigma_m^2 rval              (QCalculator.h:158-229)
//
// The per-dimension Jackson factors sigma (QCalculator.h:67-74), rval
// (QCalculator.h:176, half-width L instead of the hard-coded 6.62132) and the
// diagonal terms Bmm (QCalculator.h:181-192) are evaluated once on the host
// with libm, exactly as the reference does.  The N^6 main entries are
// evaluated on the device.  F_alpha's special branches (QCalculator.h:81-103)
// are selected on the integer squared lengths |l+m|^2 and |l-m|^2, which is
// exactly the reference's floating-point test on lattice arguments
// (lambda > 1e-10, and |xi - eta| < 1e-10 only when the integers are equal).
// Multiplies and adds are kept separately rounded, in the reference's order.
#include <cmath>
#include <vector>

#include "edb_internal.hpp"

namespace edb {
namespace {

#pragma clang fp contract(off)

__host__ __device__ inline double funF_lattice(int alpha, long A, long Bq, double lambda) {
  double xi = sqrt((double)A) * lambda;
  double eta = sqrt((double)Bq) * lambda;
  double p = xi + eta, q = xi - eta;
  double res = 0.;
  if (alpha == 0) {
    if (Bq == 0) {
      if (A == 0) res = 1. / 3.;
      else res = (-xi * cos(xi) + sin(xi)) / (xi * xi * xi);
    } else {
      if (A == 0) res = (-eta * cos(eta) + sin(eta)) / (eta * eta * eta);
      else if (A == Bq) res = (xi - cos(xi) * sin(xi)) / (2. * xi * xi * xi);
      else res = (p * sin(q) - q * sin(p)) / (2. * xi * eta * p * q);
    }
  } else if (alpha == 1) {
    if (Bq == 0) {
      if (A == 0) res = 1. / 4.;
      else res = (-2. - (-2. + xi * xi) * cos(xi) + 2. * xi * sin(xi)) / (xi * xi * xi * xi);
    } else {
      if (A == 0)
        res = (-2. - (-2. + eta * eta) * cos(eta) + 2. * eta * sin(eta)) / (eta * eta * eta * eta);
      else if (A == Bq)
        res = -(-1. - 2. * xi * xi + cos(2. * xi) + 2. * xi * sin(2. * xi)) / (8. * xi * xi * xi * xi);
      else
        res = (q * sin(q) + cos(q)) / (2. * xi * eta * q * q) -
              (p * sin(p) + cos(p)) / (2. * xi * eta * p * p) - 2. / (p * p * q * q);
    }
  }
  return res;
}

struct BetaArgs {
  int N, n, M, alpha, filter;
  double lambda, rval;
  const double* sig;
  const double* bmm;
};

// beta(l, [k-l]_N) for centred indices of l and k; 0 when [k-l]_N is the
// Nyquist residue (even N).
__device__ inline double beta_lk(const BetaArgs& a, int l1c, int l2c, int l3c, int k1c, int k2c,
                                 int k3c) {
  int n = a.n;
  int l1 = l1c - n, l2 = l2c - n, l3 = l3c - n;
  int m1, m2, m3;
  if (!wrap_freq(a.N, k1c - l1c, &m1) || !wrap_freq(a.N, k2c - l2c, &m2) ||
      !wrap_freq(a.N, k3c - l3c, &m3))
    return 0.;
  long A = (long)(l1 + m1) * (l1 + m1) + (long)(l2 + m2) * (l2 + m2) + (long)(l3 + m3) * (l3 + m3);
  // m = -l: F(0, lambda|2l|) sigma_l^2 rval - F(lambda|2l|, 0) sigma_l^2 rval, which the
  // reference evaluates to exactly 0 (same formula and arguments, QCalculator.h:84,87);
  // Bmm comes from host libm here, so the zero is kept exact explicitly.  It is what
  // makes Q_0 == 0, i.e. exact mass conservation.
  if (A == 0) return 0.;
  long Bq = (long)(l1 - m1) * (l1 - m1) + (long)(l2 - m2) * (l2 - m2) + (long)(l3 - m3) * (l3 - m3);
  double sigl = 1., sigm = 1.;
  if (a.filter) {
    sigl = a.sig[l1c] * a.sig[l2c] * a.sig[l3c];
    sigm = a.sig[m1 + n] * a.sig[m2 + n] * a.sig[m3 + n];
  }
  long mc = ((long)(m1 + n) * a.M + (m2 + n)) * a.M + (m3 + n);
  return funF_lattice(a.alpha, A, Bq, a.lambda) * sigl * sigm * a.rval - a.bmm[mc];
}

__global__ void genb_ref64_kernel(BetaArgs a, double* __restrict__ B, long NQ) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NQ * NQ) return;
  long lc = idx / NQ, kc = idx % NQ;
  int M = a.M;
  B[idx] = beta_lk(a, (int)(lc / (M * M)), (int)((lc / M) % M), (int)(lc % M), (int)(kc / (M * M)),
                   (int)((kc / M) % M), (int)(kc % M));
}

// Plain-kernel layout: the reference's row-major [l][k] with each row padded
// to ntile*64 entries (zeros past M^3), k = tile*64 + lane.
template <typename T>
__global__ void genb_plain_kernel(BetaArgs a, T* __restrict__ B, long NQ, int ntile) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long kp = (long)ntile * 64;
  long total = kp * NQ;
  if (idx >= total) return;
  long lc = idx / kp;
  long kc = idx % kp;
  int M = a.M;
  double v = 0.;
  if (kc < NQ)
    v = beta_lk(a, (int)(lc / (M * M)), (int)((lc / M) % M), (int)(lc % M), (int)(kc / (M * M)),
                (int)((kc / M) % M), (int)(kc % M));
  B[idx] = (T)v;  // FP32 layout: the FP64 value rounded once
}

// Paired layout: block b holds an M x M table [l3][k3] for the scheduled
// l-row (l1,l2) of one k-row: beta(l,m) + beta(m,l) with m = (m1, m2,
// [k3-l3]), or beta(l,m) alone on the row where (l1,l2) == (m1,m2).  The sum
// is formed in FP64 and rounded once to FP32.
__global__ void genb_pair32_kernel(BetaArgs a, const uint32_t* __restrict__ sched, long nblocks,
                                   float* __restrict__ B) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int M = a.M;
  long MM = (long)M * M;
  if (idx >= nblocks * MM) return;
  long blk = idx / MM;
  int within = (int)(idx % MM);
  int l3c = within / M, k3c = within % M;
  uint32_t e = sched[blk];
  int l1c = e & 255, l2c = (e >> 8) & 255, m1c = (e >> 16) & 255, m2c = (e >> 24) & 255;
  int n = a.n, k1, k2;
  wrap_freq(a.N, (l1c - n) + (m1c - n), &k1);
  wrap_freq(a.N, (l2c - n) + (m2c - n), &k2);
  int k1c = k1 + n, k2c = k2 + n;
  int m3;
  float out = 0.f;
  if (wrap_freq(a.N, k3c - l3c, &m3)) {
    double v = beta_lk(a, l1c, l2c, l3c, k1c, k2c, k3c);
    if (!(l1c == m1c && l2c == m2c)) v = v + beta_lk(a, m1c, m2c, m3 + n, k1c, k2c, k3c);
    out = (float)v;
  }
  B[idx] = out;
}

// Compressed beta.  beta(l, m) depends on (l, m) only through
// A = |l+m|^2, Bq = |l-m|^2, sigma_l, sigma_m and Bmm[m]; A + Bq = 2|l|^2 +
// 2|m|^2 is even, so F is tabulated over A in [0, R) and Bq/2 in [0, Rh),
// R = 12n^2 + 1, Rh = 6n^2 + 1: T[A * Rh + Bq/2] = F_alpha(lambda sqrt(A),
// lambda sqrt(Bq)), the same device evaluation beta_lk makes, so beta formed
// from the tables equals the stored beta bit for bit.  Entries whose Bq would
// exceed R - 1 are never read and hold 0.
__global__ void genb_ftable_kernel(int alpha, double lambda, long R, long Rh, double* __restrict__ T) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= R * Rh) return;
  long A = idx / Rh;
  long Bq = 2 * (idx % Rh) + (A & 1);
  T[idx] = Bq < R ? funF_lattice(alpha, A, Bq, lambda) : 0.;
}

// Per centred mode c: (sigma_c1 sigma_c2 sigma_c3, Bmm[c]), the product in
// beta_lk's order; (1, Bmm[c]) with the filter off.
__global__ void genb_sigbmm_kernel(BetaArgs a, double2* __restrict__ SB, long NQ) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= NQ) return;
  int M = a.M;
  int c1 = (int)(c / ((long)M * M)), c2 = (int)((c / M) % M), c3 = (int)(c % M);
  double sg = 1.;
  if (a.filter) sg = a.sig[c1] * a.sig[c2] * a.sig[c3];
  SB[c] = make_double2(sg, a.bmm[c]);
}

BetaArgs make_args(const BetaTables& t) {
  Geometry g(t.N);
  BetaArgs a;
  a.N = t.N; a.n = g.n; a.M = g.M; a.alpha = t.alpha; a.filter = t.filter;
  a.lambda = t.lambda; a.rval = t.rval; a.sig = t.d_sig; a.bmm = t.d_bmm;
  return a;
}

double filtering_host(int N, int beta) {  // QCalculator.h:67-74
  int n = (N - 1) / 2;
  double tmp = (double)(n + 1);
  double PI = 4. * atan(1.);
  double b = fabs((double)beta);
  return ((tmp - b) * cos(PI * b / tmp) + sin(PI * b / tmp) / tan(PI / tmp)) / tmp;
}

unsigned blocks_for(long total, int bs) { return (unsigned)((total + bs - 1) / bs); }

}  // namespace

int beta_tables_init(BetaTables* t, int N, int alpha, double L, int filter) {
  Geometry g(N);
  t->N = N; t->alpha = alpha; t->L = L; t->filter = filter;
  t->lambda = (2. / (3. + sqrt(2.))) * M_PI;                                    // :163
  t->rval = (4. * M_PI) * pow(2. * t->lambda / M_PI * L / N, (double)(3 + alpha));  // :176
  std::vector<double> sig(g.M), bmm(g.NQ);
  for (int c = 0; c < g.M; ++c) sig[c] = filter ? filtering_host(N, c - g.n) : 1.;
  long i = 0;
  for (int k1 = -g.n; k1 <= g.n; ++k1)  // Bmm, QCalculator.h:181-192
    for (int k2 = -g.n; k2 <= g.n; ++k2)
      for (int k3 = -g.n; k3 <= g.n; ++k3) {
        double fil = filter ? sig[k1 + g.n] * sig[k2 + g.n] * sig[k3 + g.n] : 1.;
        long A = 4L * (k1 * k1 + k2 * k2 + k3 * k3);
        bmm[i++] = funF_lattice(alpha, A, 0, t->lambda) * fil * fil * t->rval;
      }
  t->d_sig = t->d_bmm = nullptr;
  hipError_t e = hipMalloc(&t->d_sig, sizeof(double) * g.M);
  if (e == hipSuccess) e = hipMalloc(&t->d_bmm, sizeof(double) * g.NQ);
  if (e == hipSuccess) e = hipMemcpy(t->d_sig, sig.data(), sizeof(double) * g.M, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->d_bmm, bmm.data(), sizeof(double) * g.NQ, hipMemcpyHostToDevice);
  if (e != hipSuccess) {  // nothing half-built is kept
    set_error(std::string("beta tables: ") + hipGetErrorString(e));
    (void)hipGetLastError();
    beta_tables_free(t);
    return e == hipErrorOutOfMemory ? kNoMemory : kHipError;
  }
  return kOk;
}

void beta_tables_free(BetaTables* t) {
  if (t->d_sig) (void)hipFree(t->d_sig);
  if (t->d_bmm) (void)hipFree(t->d_bmm);
  t->d_sig = t->d_bmm = nullptr;
}

int launch_genb_ref64(const BetaTables& t, double* d_B, hipStream_t s) {
  Geometry g(t.N);
  long total = g.NQ * g.NQ;
  hipLaunchKernelGGL(genb_ref64_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, make_args(t),
                     d_B, g.NQ);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_genb_plain64(const BetaTables& t, double* d_B, int ntile, hipStream_t s) {
  Geometry g(t.N);
  long total = (long)ntile * g.NQ * 64;
  hipLaunchKernelGGL(genb_plain_kernel<double>, dim3(blocks_for(total, 256)), dim3(256), 0, s,
                     make_args(t), d_B, g.NQ, ntile);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_genb_plain32(const BetaTables& t, float* d_B, int ntile, hipStream_t s) {
  Geometry g(t.N);
  long total = (long)ntile * g.NQ * 64;
  hipLaunchKernelGGL(genb_plain_kernel<float>, dim3(blocks_for(total, 256)), dim3(256), 0, s,
                     make_args(t), d_B, g.NQ, ntile);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_genb_pair32(const BetaTables& t, const int* d_rowinfo, const uint32_t* d_sched, int nrows,
                       long nblocks, float* d_B, hipStream_t s) {
  (void)d_rowinfo;
  (void)nrows;
  Geometry g(t.N);
  long total = nblocks * g.M * g.M;
  hipLaunchKernelGGL(genb_pair32_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s,
                     make_args(t), d_sched, nblocks, d_B);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

void compressed_beta_dims(int N, long* R, long* Rh) {
  Geometry g(N);
  *R = 12L * g.n * g.n + 1;
  *Rh = 6L * g.n * g.n + 1;
}

int launch_genb_compressed(const BetaTables& t, double* d_T, double* d_SB, hipStream_t s) {
  Geometry g(t.N);
  long R, Rh;
  compressed_beta_dims(t.N, &R, &Rh);
  hipLaunchKernelGGL(genb_ftable_kernel, dim3(blocks_for(R * Rh, 256)), dim3(256), 0, s, t.alpha, t.lambda, R,
                     Rh, d_T);
  EDB_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(genb_sigbmm_kernel, dim3(blocks_for(g.NQ, 256)), dim3(256), 0, s, make_args(t),
                     (double2*)d_SB, g.NQ);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace edb
