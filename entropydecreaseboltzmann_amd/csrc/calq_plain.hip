// calq_plain.hip -- the plain single-instance FP64 evaluation of
//   Q_k = sum_l (f_l * f_[k-l]) * beta(l, [k-l])      (QCalculator.h:231-257)
//
// This is the reference algorithm on the GPU with no symmetry reduction: every
// one of the M^6 FP64 beta entries is streamed once from HBM, so the kernel
// is HBM-bound (8 bytes and 10 flops per term).  The same kernel also streams
// an FP32 copy of beta (4 bytes per term; products and sums stay FP64).
//
// Layout.  beta is the reference's row-major [l][k] (rows padded to a multiple
// of 64): a wave owns 64 consecutive centred k (one per lane) and walks a slab
// of l, so each l step is one coalesced 512-byte load, and the waves of a slab
// sweep one contiguous front of memory together (DRAM row locality).  beta is
// read once, with non-temporal loads, so it does not evict the gather table.  f_l is wave-uniform (scalar load).  f_[k-l] is
// gathered from a duplicated table D over differences d = k-l in
// [-(M-1), M-1]^3 (the role of QCalculator::transf, QCalculator.h:108-123),
// D[d] = f_[d]_N or 0 on the Nyquist residue, so the lane's address is its
// own base minus a uniform offset per l.  Slabs of l are summed afterwards in
// a fixed order (deterministic).
#include "edb_internal.hpp"

namespace edb {
namespace {

__global__ void build_dup_kernel(int N, const double2* __restrict__ fc, int stride,
                                 double2* __restrict__ dup) {
  int n = (N - 1) / 2, M = 2 * n + 1, W = 2 * M - 1;
  long total = (long)W * W * W;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  int d1 = (int)(idx / ((long)W * W)) - (M - 1);
  int d2 = (int)((idx / W) % W) - (M - 1);
  int d3 = (int)(idx % W) - (M - 1);
  int w1, w2, w3;
  double2 v = make_double2(0., 0.);
  if (wrap_freq(N, d1, &w1) && wrap_freq(N, d2, &w2) && wrap_freq(N, d3, &w3))
    v = fc[(((long)(w1 + n) * M + (w2 + n)) * M + (w3 + n)) * stride];
  dup[idx] = v;
}

// One term of the sum: q += (f_l * f_[k-l]) * beta, with the FMAs written out
// so that every kernel below rounds identically.
__device__ __forceinline__ void accum_term(double2 a, double2 c, double b, double& qr, double& qi) {
#pragma clang fp contract(off)
  const double pr = fma(a.x, c.x, -(a.y * c.y));
  const double pi = fma(a.x, c.y, a.y * c.x);
  qr = fma(pr, b, qr);
  qi = fma(pi, b, qi);
}

template <int UNROLL, typename T>
__global__ __launch_bounds__(64) void calq_plain64_kernel(int M, long NQ, const T* __restrict__ B,
                                                          const double2* __restrict__ dup,
                                                          const double2* __restrict__ fc, int stride,
                                                          double2* __restrict__ part, int nslab) {
  const int W = 2 * M - 1;
  const int lane = threadIdx.x;
  const int tile = blockIdx.x;
  const int slab = blockIdx.y;
  long kc = (long)tile * 64 + lane;
  long kcc = kc < NQ ? kc : NQ - 1;  // padded lanes read a valid address; their beta is 0
  int k1 = (int)(kcc / ((long)M * M)), k2 = (int)((kcc / M) % M), k3 = (int)(kcc % M);
  long dbase = ((long)(k1 + M - 1) * W + (k2 + M - 1)) * W + (k3 + M - 1);
  long l0 = NQ * slab / nslab, l1e = NQ * (slab + 1) / nslab;
  const long kp = (long)gridDim.x * 64;  // padded row length
  const T* bt = B + (long)tile * 64 + lane;
  double qr = 0., qi = 0.;
  int a1 = (int)(l0 / ((long)M * M)), a2 = (int)((l0 / M) % M), a3 = (int)(l0 % M);
  long l = l0;
  for (; l + UNROLL <= l1e; l += UNROLL) {
    double b[UNROLL];
    double2 c[UNROLL], a[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      long lin = ((long)a1 * W + a2) * W + a3;
      b[u] = (double)__builtin_nontemporal_load(bt + (l + u) * kp);
      c[u] = dup[dbase - lin];
      a[u] = fc[(l + u) * stride];
      if (++a3 == M) { a3 = 0; if (++a2 == M) { a2 = 0; ++a1; } }
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) accum_term(a[u], c[u], b[u], qr, qi);
  }
  for (; l < l1e; ++l) {
    long lin = ((long)a1 * W + a2) * W + a3;
    double b = (double)__builtin_nontemporal_load(bt + l * kp);
    double2 c = dup[dbase - lin];
    double2 a = fc[l * stride];
    accum_term(a, c, b, qr, qi);
    if (++a3 == M) { a3 = 0; if (++a2 == M) { a2 = 0; ++a1; } }
  }
  if (kc < NQ) part[(long)slab * NQ + kc] = make_double2(qr, qi);
}

// Frequency of a centred-index difference d in [-(M-1), M-1]: wrap_freq
// without the division (valid = false on the Nyquist residue of even N,
// where w is set to 0 so that every table index stays in range).
__device__ __forceinline__ int wrap_small(int N, int n, int d, bool* ok) {
  int w = d > n ? d - N : (d < -n ? d + N : d);
  *ok = w >= -n && w <= n;
  return *ok ? w : 0;
}

// The plain evaluation with beta formed on the fly from the compressed tables
// (genb.hip): beta(l, m) = T[|l+m|^2 Rh + |l-m|^2 / 2] sigma_l sigma_m rval -
// Bmm[m], in the stored kernel's order, so the result is the stored FP64
// kernel's bit for bit.  Nothing of size M^6 exists: the reads per term are
// the F entry (L2/MALL-resident table), the (sigma, Bmm) pair of m and the
// f_[k-l] gather.  The (l1, l2) parts of the index arithmetic are kept per
// l-row; each term does the third dimension only.
template <int UNROLL>
__global__ __launch_bounds__(64) void calq_plain_otf_kernel(int N, int M, long NQ, double rval,
                                                            const double* __restrict__ T, long Rh,
                                                            const double2* __restrict__ SB,
                                                            const double2* __restrict__ dup,
                                                            const double2* __restrict__ fc, int stride,
                                                            double2* __restrict__ part, int nslab) {
  const int n = (N - 1) / 2, W = 2 * M - 1;
  const int lane = threadIdx.x;
  const int tile = blockIdx.x;
  const int slab = blockIdx.y;
  long kc = (long)tile * 64 + lane;
  long kcc = kc < NQ ? kc : NQ - 1;  // padded lanes compute a valid term that is never written
  int k1 = (int)(kcc / ((long)M * M)), k2 = (int)((kcc / M) % M), k3 = (int)(kcc % M);
  long dbase = ((long)(k1 + M - 1) * W + (k2 + M - 1)) * W + (k3 + M - 1);
  long l0 = NQ * slab / nslab, l1e = NQ * (slab + 1) / nslab;
  int a1 = (int)(l0 / ((long)M * M)), a2 = (int)((l0 / M) % M), a3 = (int)(l0 % M);
  // per l-row (a1, a2): validity and the (1, 2) parts of |l+m|^2, |l-m|^2, m
  bool rowok;
  long A12, B12, mc12;
  auto row = [&]() {
    bool v1, v2;
    const int m1 = wrap_small(N, n, k1 - a1, &v1), m2 = wrap_small(N, n, k2 - a2, &v2);
    const int l1 = a1 - n, l2 = a2 - n;
    rowok = v1 && v2;
    A12 = (long)(l1 + m1) * (l1 + m1) + (long)(l2 + m2) * (l2 + m2);
    B12 = (long)(l1 - m1) * (l1 - m1) + (long)(l2 - m2) * (l2 - m2);
    mc12 = ((long)(m1 + n) * M + (m2 + n)) * M + n;
  };
  auto beta = [&](long l) -> double {
#pragma clang fp contract(off)
    bool v3;
    const int m3 = wrap_small(N, n, k3 - a3, &v3);
    const int l3 = a3 - n;
    const long A = A12 + (long)(l3 + m3) * (l3 + m3);
    const long Bq = B12 + (long)(l3 - m3) * (l3 - m3);
    const double F = T[A * Rh + (Bq >> 1)];
    const double2 sb = SB[mc12 + m3];
    const double v = F * SB[l].x * sb.x * rval - sb.y;
    return (rowok && v3 && A != 0) ? v : 0.;
  };
  row();
  double qr = 0., qi = 0.;
  long l = l0;
  for (; l + UNROLL <= l1e; l += UNROLL) {
    double b[UNROLL];
    double2 c[UNROLL], a[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      long lin = ((long)a1 * W + a2) * W + a3;
      c[u] = dup[dbase - lin];
      a[u] = fc[(l + u) * stride];
      b[u] = beta(l + u);
      if (++a3 == M) { a3 = 0; if (++a2 == M) { a2 = 0; ++a1; } row(); }
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) accum_term(a[u], c[u], b[u], qr, qi);
  }
  for (; l < l1e; ++l) {
    long lin = ((long)a1 * W + a2) * W + a3;
    double2 c = dup[dbase - lin];
    double2 a = fc[l * stride];
    double b = beta(l);
    accum_term(a, c, b, qr, qi);
    if (++a3 == M) { a3 = 0; if (++a2 == M) { a2 = 0; ++a1; } row(); }
  }
  if (kc < NQ) part[(long)slab * NQ + kc] = make_double2(qr, qi);
}

__global__ void reduce_plain_kernel(long NQ, const double2* __restrict__ part, int nslab,
                                    double2* __restrict__ q, int stride) {
  long kc = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (kc >= NQ) return;
  double qr = 0., qi = 0.;
  for (int s = 0; s < nslab; ++s) {
    double2 v = part[(long)s * NQ + kc];
    qr += v.x;
    qi += v.y;
  }
  q[kc * stride] = make_double2(qr, qi);
}

}  // namespace

int launch_build_dup(int N, const double* d_fc, int fc_stride, double* d_dup, hipStream_t s) {
  Geometry g(N);
  long W = 2 * g.M - 1, total = W * W * W;
  hipLaunchKernelGGL(build_dup_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, N,
                     (const double2*)d_fc, fc_stride, (double2*)d_dup);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_calq_plain64(int N, const double* d_B, const double* d_dup, const double* d_fc,
                        int fc_stride, double* d_part, int nslab, hipStream_t s) {
  Geometry g(N);
  EDB_REQUIRE(nslab >= 1 && nslab <= 4096 && nslab <= g.NQ, "calq_plain64: bad slab count");
  int ntile = (int)((g.NQ + 63) / 64);
  hipLaunchKernelGGL((calq_plain64_kernel<8, double>), dim3(ntile, nslab), dim3(64), 0, s, g.M, g.NQ, d_B,
                     (const double2*)d_dup, (const double2*)d_fc, fc_stride, (double2*)d_part, nslab);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

// Same evaluation with the FP32 copy of beta (products and sums stay FP64):
// half the bytes of the HBM stream.
int launch_calq_plain32(int N, const float* d_B, const double* d_dup, const double* d_fc, int fc_stride,
                        double* d_part, int nslab, hipStream_t s) {
  Geometry g(N);
  EDB_REQUIRE(nslab >= 1 && nslab <= 4096 && nslab <= g.NQ, "calq_plain32: bad slab count");
  int ntile = (int)((g.NQ + 63) / 64);
  hipLaunchKernelGGL((calq_plain64_kernel<8, float>), dim3(ntile, nslab), dim3(64), 0, s, g.M, g.NQ, d_B,
                     (const double2*)d_dup, (const double2*)d_fc, fc_stride, (double2*)d_part, nslab);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_calq_plain_otf(int N, double rval, const double* d_T, const double* d_SB, const double* d_dup,
                          const double* d_fc, int fc_stride, double* d_part, int nslab, hipStream_t s) {
  Geometry g(N);
  EDB_REQUIRE(nslab >= 1 && nslab <= 4096 && nslab <= g.NQ, "calq_plain_otf: bad slab count");
  long R, Rh;
  compressed_beta_dims(N, &R, &Rh);
  int ntile = (int)((g.NQ + 63) / 64);
  hipLaunchKernelGGL((calq_plain_otf_kernel<8>), dim3(ntile, nslab), dim3(64), 0, s, N, g.M, g.NQ, rval, d_T, Rh,
                     (const double2*)d_SB, (const double2*)d_dup, (const double2*)d_fc, fc_stride,
                     (double2*)d_part, nslab);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

int launch_reduce_plain(int N, const double* d_part, int nslab, double* d_q, int q_stride,
                        hipStream_t s) {
  Geometry g(N);
  hipLaunchKernelGGL(reduce_plain_kernel, dim3((unsigned)((g.NQ + 255) / 256)), dim3(256), 0, s,
                     g.NQ, (const double2*)d_part, nslab, (double2*)d_q, q_stride);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace edb
