// ops.hip -- the O(N^3) kernels around the collision operator: RK stage
// combinations, packing of stage inputs for the Q kernels, the entropy of
// BKW3D, the EntropyFix bisection, FP64 moments and BKW error norms.
//
// All state arrays are instance-minor: [grid point or mode][ld] with one lane
// per instance, so every elementwise pass is coalesced across instances.
// Per-instance reductions use one workgroup per instance and a fixed-shape
// tree, so results are deterministic.
#include <algorithm>
#include <cmath>

#include "edb_internal.hpp"
#include "ops.hpp"

namespace edb {
namespace {

// Galerkin centred index of FFTW grid point j (or -1: Nyquist plane, even N).
__device__ inline long centred_of(int N, long j) {
  int n = (N - 1) / 2, M = 2 * n + 1;
  int j1 = (int)(j / ((long)N * N)), j2 = (int)((j / N) % N), j3 = (int)(j % N);
  auto c = [&](int x) { return x <= n ? x + n : (x >= N - n ? x - N + n : -1); };
  int c1 = c(j1), c2 = c(j2), c3 = c(j3);
  if (c1 < 0 || c2 < 0 || c3 < 0) return -1;
  return ((long)c1 * M + c2) * M + c3;
}

__device__ inline double2 qget(const double2* Q, long c, long ld, long col) {
  return c >= 0 ? Q[c * ld + col] : make_double2(0., 0.);
}

// Stage combinations in the reference's operation order (BKW.h).  Each is
// applied on all N^3 grid points; Q is 0 off the Galerkin set.
__global__ void combine_kernel(int op, int N, int batch, int ld, double2* out, const double2* X,
                               const double2* Y, const double2* Q0, const double2* Q1,
                               const double2* Q2, const double2* Q3, double c0, double c1, double c2) {
  long N3 = (long)N * N * N;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N3 * ld) return;
  long j = idx / ld, col = idx % ld;
  if (col >= batch) return;
  long c = centred_of(N, j);
  double2 x = X[idx];
  double2 r;
  switch (op) {
    case kOpAxpy: {  // out = X + c0*Q0   (BKW.h:222, 343, 346, 418, 484, 487, 490)
      double2 q = qget(Q0, c, ld, col);
      r = make_double2(x.x + c0 * q.x, x.y + c0 * q.y);
      break;
    }
    case kOpRK4: {  // out = X + c0*(Q0 + 2*Q1 + 2*Q2 + Q3)   (BKW.h:494)
      double2 a = qget(Q0, c, ld, col), b = qget(Q1, c, ld, col), d = qget(Q2, c, ld, col),
              e = qget(Q3, c, ld, col);
      double sr = a.x + 2. * b.x + 2. * d.x + e.x;
      double si = a.y + 2. * b.y + 2. * d.y + e.y;
      r = make_double2(x.x + c0 * sr, x.y + c0 * si);
      break;
    }
    case kOpSSP3b: {  // out = 0.75*X + 0.25*Y + c0*Q0   (BKW.h:421)
      double2 y = Y[idx], q = qget(Q0, c, ld, col);
      r = make_double2(0.75 * x.x + 0.25 * y.x + c0 * q.x, 0.75 * x.y + 0.25 * y.y + c0 * q.y);
      break;
    }
    case kOpSSP3c: {  // out = X/3 + 2*Y/3 + c0*Q0   (BKW.h:424)
      double2 y = Y[idx], q = qget(Q0, c, ld, col);
      r = make_double2(x.x / 3. + 2. * y.x / 3. + c0 * q.x, x.y / 3. + 2. * y.y / 3. + c0 * q.y);
      break;
    }
    case kOpDufort: {  // out = c0*(X + c1*Q0 + c2*(X - Y))   (BKW.h:598)
      double2 y = Y[idx], q = qget(Q0, c, ld, col);
      r = make_double2(c0 * (x.x + c1 * q.x + c2 * (x.x - y.x)),
                       c0 * (x.y + c1 * q.y + c2 * (x.y - y.y)));
      break;
    }
    default:  // copy
      r = x;
  }
  out[idx] = r;
}

// Stage input for the Q kernels: the Galerkin modes of X (+ c0*Q0) in centred
// order, FP64 (plain kernel) or rounded to FP32 (batched kernel).
template <typename T>
__global__ void pack_kernel(int N, int batch, int ld, T* out, const double2* X, const double2* Q0,
                            double c0) {
  int n = (N - 1) / 2, M = 2 * n + 1;
  long NQ = (long)M * M * M;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NQ * ld) return;
  long c = idx / ld, col = idx % ld;
  int c1 = (int)(c / ((long)M * M)), c2 = (int)((c / M) % M), c3 = (int)(c % M);
  long j = ((long)fftw_index(N, c1 - n) * N + fftw_index(N, c2 - n)) * N + fftw_index(N, c3 - n);
  double2 v = make_double2(0., 0.);
  if (col < batch) {
    v = X[j * ld + col];
    if (Q0) {
      double2 q = Q0[idx];
      v = make_double2(v.x + c0 * q.x, v.y + c0 * q.y);
    }
  }
  if constexpr (sizeof(T) == 8) out[idx] = make_float2((float)v.x, (float)v.y);
  else out[idx] = v;
}

// Centred-order Q of one column to an FFTW-order grid (convertQ,
// QCalculator.h:125-156); Nyquist points get 0.
__global__ void unpack_kernel(int N, int batch, int ld, double2* out, const double2* Q) {
  long N3 = (long)N * N * N;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N3 * ld) return;
  long j = idx / ld, col = idx % ld;
  long c = centred_of(N, j);
  out[idx] = (col < batch) ? qget(Q, c, ld, col) : make_double2(0., 0.);
}

// Column-major transposes between the caller's [batch][n] host order and the
// device's [n][ld] instance-minor order.
__global__ void gather_cols_kernel(long n, int batch, int ld, const double2* __restrict__ src,
                                   double2* __restrict__ dst) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over n*ld, dst index
  if (idx >= n * ld) return;
  long j = idx / ld, col = idx % ld;
  dst[idx] = col < batch ? src[col * n + j] : make_double2(0., 0.);
}

__global__ void scatter_cols_kernel(long n, int batch, int ld, const double2* __restrict__ src,
                                    double2* __restrict__ dst) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over batch*n, dst index
  if (idx >= n * batch) return;
  long col = idx / n, j = idx % n;
  dst[idx] = src[j * ld + col];
}

__global__ void real_to_complex_cols_kernel(long n, int batch, int ld, const double* __restrict__ src,
                                            double2* __restrict__ dst) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * ld) return;
  long j = idx / ld, col = idx % ld;
  dst[idx] = make_double2(col < batch ? src[col * n + j] : 0., 0.);
}

__global__ void complex_to_real_cols_kernel(long n, int batch, int ld, const double2* __restrict__ src,
                                            double* __restrict__ dst) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * batch) return;
  long col = idx / n, j = idx % n;
  dst[idx] = src[j * ld + col].x;
}

__global__ void fill_kernel(double* p, long n, double v) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) p[idx] = v;
}

// Copy column col of src to dst where flag[col] != 0.
__global__ void select_cols_kernel(long n, int batch, int ld, const int* flag, const double2* src,
                                   double2* dst) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * ld) return;
  long col = idx % ld;
  if (col < batch && flag[col]) dst[idx] = src[idx];
}

constexpr int kRedThreads = 512;

__device__ inline double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// BKW3D::flogf, BKW.h:153-161.
__device__ inline double flogf_dev(double f) { return fabs(f) < 1.0e-6 ? 0. : f * log(f) - f; }

// EntropyFix::FixSol(Nf, fin, Aim), EntropyFix.h:211-253, by one block on
// the contiguous vector f[0..N3).  Bisection on beta in [0,1] for
// f_beta = beta*avg + (1-beta) f until the width is <= eps; each probe is
// H(f_beta) = (sum_{f_beta > eps} f_beta log f_beta)/Nf - avg (EntropyCal,
// EntropyFix.h:124-161).  Early exit when Aim-eps < H < Aim, else
// f <- f_{betaR}.  Every thread sees the same block sums, so the loop is
// uniform.  f is rewritten in place.
//
// The blend and the final update round each product and sum separately, as
// the reference's loops do (EntropyFix.h:219-222, 249-252).  With `term`
// non-null every entropy sum runs in the reference's index order (the terms
// are formed in parallel into term[0..N3), then one thread adds them up), so
// the bisection takes the reference's decisions wherever its margin exceeds
// the difference of the device's log from the host's; with `term` null the
// sums are parallel block reductions.
__device__ double seq_sum(const double* term, long n, double* sh) {
  __syncthreads();  // every thread's terms written
  if (threadIdx.x == 0) {
    double s = 0.;
    for (long i = 0; i < n; ++i) s += term[i];
    sh[0] = s;
  }
  __syncthreads();
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ inline double blend(double betaFbar, double betaMid, double f) {
#pragma clang fp contract(off)
  return betaFbar + (1. - betaMid) * f;
}

__device__ inline double blend_final(double betaR, double betaFbar, double f) {
#pragma clang fp contract(off)
  return f * (1.0 - betaR) + betaFbar;
}

__device__ inline double xlogx_above(double f, double eps) { return f > eps ? f * log(f) : 0.; }

__device__ void fixsol_block(double* f, long N3, double aim, double eps, double avg, double* sh,
                             double* term = nullptr) {
  const int t = threadIdx.x;
  double betaL = 0.0, betaR = 1.0;
  while ((betaR - betaL) > eps) {
    const double betaMid = 0.5 * (betaL + betaR);
    const double betaFbar = betaMid * avg;
    double sum;
    if (term) {
      for (long g = t; g < N3; g += kRedThreads) term[g] = xlogx_above(blend(betaFbar, betaMid, f[g]), eps);
      sum = seq_sum(term, N3, sh);
    } else {
      double s = 0.;
#pragma unroll 4
      for (long g = t; g < N3; g += kRedThreads) s += xlogx_above(blend(betaFbar, betaMid, f[g]), eps);
      sum = block_sum(s, sh);
    }
    const double entMid = sum / (double)N3 - avg;
    if ((entMid < aim) && ((aim - entMid) < eps)) {
      for (long g = t; g < N3; g += kRedThreads) f[g] = blend(betaFbar, betaMid, f[g]);
      __syncthreads();
      return;
    }
    if (entMid < aim) betaR = betaMid;
    else betaL = betaMid;
  }
  const double betaFbar = betaR * avg;
  for (long g = t; g < N3; g += kRedThreads) f[g] = blend_final(betaR, betaFbar, f[g]);
  __syncthreads();
}

// Coalesced first pass: lane = instance, wave = (instance group, grid chunk);
// each lane sums flogf over its chunk of grid points.
__global__ __launch_bounds__(256) void entropy_part_kernel(long N3, int ld, int groups, int chunks,
                                                           const double2* __restrict__ F,
                                                           double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const long wid = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wid >= (long)groups * chunks) return;
  const int c = (int)(wid / groups), g = (int)(wid % groups);
  const long col = (long)g * 64 + lane;
  const long per = (N3 + chunks - 1) / chunks;
  const long g0 = (long)c * per, g1 = g0 + per < N3 ? g0 + per : N3;
  double acc = 0.;
  for (long q = g0; q < g1; ++q) acc += flogf_dev(F[q * ld + col].x);
  part[(long)c * ld + col] = acc;
}

// Per instance (one 512-thread block each): calEntropy of Re F
// (BKW.h:140-152); if eligible and H - H0 > 1e-10, FixSol on Re F with target
// H0/(2R)^3 (the drivers' call, BKW.h:229-233, 502-506; the average cached per
// instance at its first fix, EntropyFix.h:225), then calEntropy of the fixed
// F.  H is summed from the first pass's partials; Re F is staged in the
// instance's contiguous work row W only when the fix runs.  Mode 2 writes
// the fixed F back (imaginary part 0) and flags the instance for the forward
// transform.  Records H, Hfix and whether a fix ran.
__global__ __launch_bounds__(512) void entropy_fix_kernel(EntropyFixArgs A, int chunks) {
  __shared__ double sh[kRedThreads];
  const int inst = blockIdx.x;
  const int t = threadIdx.x;
  const long N3 = A.N3;
  const int ld = A.ld;
  double* w = A.W + (long)inst * N3;
  double part = 0.;
  for (int c = t; c < chunks; c += kRedThreads) part += A.part[(long)c * ld + inst];
  const double H = block_sum(part, sh) * A.vol_over_N3;
  const double H0 = A.H0[inst];
  double Hfix = H;
  // mean of Re F: the k = 0 coefficient of the state over N^3
  const double mean = A.Fh ? A.Fh[inst].x / (double)N3 : 0.;
  double mean_fix = mean;
  int fixed = 0;
  const bool elig = A.eligible && (!A.steps_done || A.steps_done[inst] > 0);
  if (elig && A.fix_mode != 0 && H - H0 > 1.0e-10) {
    for (long g = t; g < N3; g += kRedThreads) w[g] = A.F[g * ld + inst].x;
    __syncthreads();
    double avg = A.avg[inst];
    if (!A.avg_flag[inst]) {  // setAverage(Nf, fin), EntropyFix.h:225
      double s = 0.;
      for (long g = t; g < N3; g += kRedThreads) s += w[g];
      avg = block_sum(s, sh) / (double)N3;
      if (t == 0) {
        A.avg[inst] = avg;
        A.avg_flag[inst] = 1;
      }
    }
    fixsol_block(w, N3, H0 / A.vol, A.eps, avg, sh);
    double s2 = 0., s3 = 0.;
    for (long g = t; g < N3; g += kRedThreads) {
      s2 += flogf_dev(w[g]);
      s3 += w[g];
    }
    Hfix = block_sum(s2, sh) * A.vol_over_N3;
    mean_fix = block_sum(s3, sh) / (double)N3;
    fixed = 1;
    // the driver's real-space F holds the fixed values in both modes; only
    // mode 2 feeds them back into the spectral state
    for (long g = t; g < N3; g += kRedThreads) A.F[g * ld + inst] = make_double2(w[g], 0.);
  }
  if (t == 0) {
    A.H0[inst] = Hfix;
    if (A.fixed) A.fixed[inst] = (A.fix_mode == 2) ? fixed : 0;
    const int rec = A.rec_ptr ? *A.rec_ptr : A.rec;
    if (A.hist_H && rec < A.hist_cap) {
      A.hist_H[(long)rec * ld + inst] = H;
      A.hist_Hfix[(long)rec * ld + inst] = Hfix;
      A.hist_fixed[(long)rec * ld + inst] = fixed;
      if (A.hist_mean) {
        A.hist_mean[(long)rec * ld + inst] = mean;
        A.hist_mean_fix[(long)rec * ld + inst] = mean_fix;
      }
    }
  }
}

__global__ void move_col_kernel(long n, int ld, int src, int dst, double2* a) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) a[j * ld + dst] = a[j * ld + src];
}

__global__ void move_slot_scalars_kernel(int src, int dst, double* d0, double* d1, int* i0, int* i1) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (d0) d0[dst] = d0[src];
  if (d1) d1[dst] = d1[src];
  if (i0) i0[dst] = i0[src];
  if (i1) i1[dst] = i1[src];
}

__global__ void incr_kernel(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += 1;
}

__global__ void extract_col_kernel(long n, int ld, int col, const double2* a, double* out) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = a[j * ld + col].x;
}

__global__ void reset_slots_kernel(int ld, const int* flag, double* H0, double* avg, int* avg_flag,
                                   int* steps) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld || !flag[i]) return;
  H0[i] = 1.;
  avg[i] = 1.;
  avg_flag[i] = 0;
  steps[i] = 0;
}

__global__ void stage_real_kernel(long n, int batch, int ld, const int* flag, const double* rows,
                                  double2* F) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * ld) return;
  long j = idx / ld, col = idx % ld;
  if (col < batch && flag[col]) F[idx] = make_double2(rows[col * n + j], 0.);
}

__global__ void bump_kernel(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p += 1;
}

// The EntropyFix members on one contiguous vector (one block):
// mode 0 setAverage, 1 EntropyCal, 2 FixSol (EntropyFix.h:65-253).
// st[0] = average, st[1] = flag, st[2] = result entropy.
// The reference's EntropyFix on one vector (EntropyFix.h:51-253), every sum in
// the reference's index order: mode 0 setAverage, 1 EntropyCal (setting the
// average on first use), 2 FixSol.  st = {average, average set, result};
// term: N3 doubles of scratch.
__global__ __launch_bounds__(512) void efix_kernel(int mode, long N3, double* f, double eps, double aim,
                                                   double* st, double* term) {
  __shared__ double sh[kRedThreads];
  const int t = threadIdx.x;
  double avg = st[0];
  if (mode == 0 || st[1] == 0.) {
    avg = seq_sum(f, N3, sh) / (double)N3;
    if (t == 0) {
      st[0] = avg;
      st[1] = 1.;
    }
  }
  if (mode == 1) {
    for (long g = t; g < N3; g += kRedThreads) term[g] = xlogx_above(f[g], eps);
    const double r = seq_sum(term, N3, sh) / (double)N3 - avg;
    if (t == 0) st[2] = r;
  } else if (mode == 2) {
    fixsol_block(f, N3, aim, eps, avg, sh, term);
  }
}

// FP64 moments of Re F on the grid x_i = -L + i*2L/N (BKW.h:121-138):
// mass, momentum (3), energy = sum |v|^2 f, each times h^3.
__global__ __launch_bounds__(512) void moments_kernel(int N, int ld, double L, const double2* F,
                                                       double* out) {
  __shared__ double sh[kRedThreads];
  const int inst = blockIdx.x;
  const long N3 = (long)N * N * N;
  const double h = 2 * L / (double)N;
  double m[5] = {0., 0., 0., 0., 0.};
  for (long g = threadIdx.x; g < N3; g += kRedThreads) {
    int i = (int)(g / ((long)N * N)), j = (int)((g / N) % N), k = (int)(g % N);
    double x = -L + i * h, y = -L + j * h, z = -L + k * h;
    double f = F[g * ld + inst].x;
    m[0] += f;
    m[1] += x * f;
    m[2] += y * f;
    m[3] += z * f;
    m[4] += (x * x + y * y + z * z) * f;
  }
  const double h3 = h * h * h;
  for (int q = 0; q < 5; ++q) {
    double s = block_sum(m[q], sh);
    if (threadIdx.x == 0) out[inst * 5 + q] = s * h3;
  }
}

// L-infinity and L1 (times h^3) differences of Re F from the BKW profile at K
// (testfft.cpp:11-16 written in K).
__global__ __launch_bounds__(512) void bkw_error_kernel(int N, int ld, double L, double K,
                                                         const double2* F, double* out) {
  __shared__ double sh[kRedThreads];
  const int inst = blockIdx.x;
  const long N3 = (long)N * N * N;
  const double h = 2 * L / (double)N;
  double linf = 0., l1 = 0.;
  for (long g = threadIdx.x; g < N3; g += kRedThreads) {
    int i = (int)(g / ((long)N * N)), j = (int)((g / N) % N), k = (int)(g % N);
    double x = -L + i * h, y = -L + j * h, z = -L + k * h;
    double v2 = x * x + y * y + z * z;
    double ex = 1. / pow(2. * M_PI * K, 1.5) * exp(-v2 / (2. * K)) *
                ((5. * K - 3.) / (2. * K) + (1. - K) / (2. * K * K) * v2);
    double d = fabs(F[g * ld + inst].x - ex);
    linf = fmax(linf, d);
    l1 += d;
  }
  // max-reduction
  sh[threadIdx.x] = linf;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  double mx = sh[0];
  __syncthreads();
  double s1 = block_sum(l1, sh);
  if (threadIdx.x == 0) {
    out[inst * 2] = mx;
    out[inst * 2 + 1] = s1 * h * h * h;
  }
}

unsigned nblk(long total, int bs) { return (unsigned)((total + bs - 1) / bs); }

}  // namespace

int launch_combine(int op, int N, int batch, int ld, double* out, const double* X, const double* Y,
                   const double* Q0, const double* Q1, const double* Q2, const double* Q3, double c0,
                   double c1, double c2, hipStream_t s) {
  long total = (long)N * N * N * ld;
  hipLaunchKernelGGL(combine_kernel, dim3(nblk(total, 256)), dim3(256), 0, s, op, N, batch, ld,
                     (double2*)out, (const double2*)X, (const double2*)Y, (const double2*)Q0,
                     (const double2*)Q1, (const double2*)Q2, (const double2*)Q3, c0, c1, c2);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_pack(int N, int batch, int ld, bool fp32, void* out, const double* X, const double* Q0,
                double c0, hipStream_t s) {
  Geometry g(N);
  long total = g.NQ * ld;
  if (fp32)
    hipLaunchKernelGGL(pack_kernel<float2>, dim3(nblk(total, 256)), dim3(256), 0, s, N, batch, ld,
                       (float2*)out, (const double2*)X, (const double2*)Q0, c0);
  else
    hipLaunchKernelGGL(pack_kernel<double2>, dim3(nblk(total, 256)), dim3(256), 0, s, N, batch, ld,
                       (double2*)out, (const double2*)X, (const double2*)Q0, c0);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_unpack(int N, int batch, int ld, double* out, const double* Q, hipStream_t s) {
  long total = (long)N * N * N * ld;
  hipLaunchKernelGGL(unpack_kernel, dim3(nblk(total, 256)), dim3(256), 0, s, N, batch, ld,
                     (double2*)out, (const double2*)Q);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_gather_cols(long n, int batch, int ld, const double* src, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(gather_cols_kernel, dim3(nblk(n * ld, 256)), dim3(256), 0, s, n, batch, ld,
                     (const double2*)src, (double2*)dst);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_scatter_cols(long n, int batch, int ld, const double* src, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(scatter_cols_kernel, dim3(nblk(n * batch, 256)), dim3(256), 0, s, n, batch, ld,
                     (const double2*)src, (double2*)dst);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_real_to_complex_cols(long n, int batch, int ld, const double* src, double* dst,
                                hipStream_t s) {
  hipLaunchKernelGGL(real_to_complex_cols_kernel, dim3(nblk(n * ld, 256)), dim3(256), 0, s, n, batch,
                     ld, src, (double2*)dst);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_complex_to_real_cols(long n, int batch, int ld, const double* src, double* dst,
                                hipStream_t s) {
  hipLaunchKernelGGL(complex_to_real_cols_kernel, dim3(nblk(n * batch, 256)), dim3(256), 0, s, n,
                     batch, ld, (const double2*)src, dst);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_fill(double* p, long n, double v, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, p, n, v);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_select_cols(long n, int batch, int ld, const int* flag, const double* src, double* dst,
                       hipStream_t s) {
  hipLaunchKernelGGL(select_cols_kernel, dim3(nblk(n * ld, 256)), dim3(256), 0, s, n, batch, ld, flag,
                     (const double2*)src, (double2*)dst);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int entropy_chunks(long N3, int ld) {
  const int groups = (ld + 63) / 64;
  long c = (4096 + groups - 1) / groups;  // about 4096 waves in the first pass
  return (int)std::max(1L, std::min(c, N3));
}

int launch_entropy_fix(const EntropyFixArgs& a, int batch, hipStream_t s) {
  EDB_REQUIRE(a.W != nullptr && a.part != nullptr && a.N3 >= 1, "entropy_fix: missing work buffer");
  EDB_REQUIRE(a.ld % 64 == 0 && batch <= a.ld, "entropy_fix: bad leading dimension");
  const int groups = (batch + 63) / 64, chunks = entropy_chunks(a.N3, a.ld);
  const long waves = (long)groups * chunks;
  hipLaunchKernelGGL(entropy_part_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, a.N3, a.ld,
                     groups, chunks, (const double2*)a.F, a.part);
  EDB_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(entropy_fix_kernel, dim3(batch), dim3(kRedThreads), 0, s, a, chunks);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_move_col(long n, int ld, int src, int dst, double* a, hipStream_t s) {
  EDB_REQUIRE(src >= 0 && dst >= 0 && src < ld && dst < ld, "move_col: column out of range");
  hipLaunchKernelGGL(move_col_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, n, ld, src, dst, (double2*)a);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_move_slot_scalars(int src, int dst, double* d0, double* d1, int* i0, int* i1, hipStream_t s) {
  hipLaunchKernelGGL(move_slot_scalars_kernel, dim3(1), dim3(64), 0, s, src, dst, d0, d1, i0, i1);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_incr(int* p, int n, hipStream_t s) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(incr_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_extract_col(long n, int ld, int col, const double* a, double* out, hipStream_t s) {
  EDB_REQUIRE(col >= 0 && col < ld, "extract_col: column out of range");
  hipLaunchKernelGGL(extract_col_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, n, ld, col,
                     (const double2*)a, out);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_reset_slots(int ld, const int* flag, double* H0, double* avg, int* avg_flag, int* steps,
                       hipStream_t s) {
  hipLaunchKernelGGL(reset_slots_kernel, dim3((ld + 255) / 256), dim3(256), 0, s, ld, flag, H0, avg,
                     avg_flag, steps);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_stage_real(long n, int batch, int ld, const int* flag, const double* rows, double* F,
                      hipStream_t s) {
  hipLaunchKernelGGL(stage_real_kernel, dim3(nblk(n * ld, 256)), dim3(256), 0, s, n, batch, ld, flag, rows,
                     (double2*)F);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_bump(int* p, hipStream_t s) {
  hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(64), 0, s, p);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_efix(int mode, long N3, double* d_f, double eps, double aim, double* d_st, double* d_term,
                hipStream_t s) {
  EDB_REQUIRE(N3 >= 1, "EntropyFix: Nf must be positive");
  EDB_REQUIRE(d_term, "EntropyFix: no term scratch");
  hipLaunchKernelGGL(efix_kernel, dim3(1), dim3(kRedThreads), 0, s, mode, N3, d_f, eps, aim, d_st, d_term);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_moments(int N, int batch, int ld, double L, const double* F, double* out, hipStream_t s) {
  hipLaunchKernelGGL(moments_kernel, dim3(batch), dim3(kRedThreads), 0, s, N, ld, L,
                     (const double2*)F, out);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_bkw_error(int N, int batch, int ld, double L, double K, const double* F, double* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(bkw_error_kernel, dim3(batch), dim3(kRedThreads), 0, s, N, ld, L, K,
                     (const double2*)F, out);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace edb
