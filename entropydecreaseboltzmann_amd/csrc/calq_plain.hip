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
    for (int u = 0; u < UNROLL; ++u) {
      double pr = a[u].x * c[u].x - a[u].y * c[u].y;
      double pi = a[u].x * c[u].y + a[u].y * c[u].x;
      qr += pr * b[u];
      qi += pi * b[u];
    }
  }
  for (; l < l1e; ++l) {
    long lin = ((long)a1 * W + a2) * W + a3;
    double b = (double)__builtin_nontemporal_load(bt + l * kp);
    double2 c = dup[dbase - lin];
    double2 a = fc[l * stride];
    double pr = a.x * c.x - a.y * c.y;
    double pi = a.x * c.y + a.y * c.x;
    qr += pr * b;
    qi += pi * b;
    if (++a3 == M) { a3 = 0; if (++a2 == M) { a2 = 0; ++a1; } }
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

int launch_reduce_plain(int N, const double* d_part, int nslab, double* d_q, int q_stride,
                        hipStream_t s) {
  Geometry g(N);
  hipLaunchKernelGGL(reduce_plain_kernel, dim3((unsigned)((g.NQ + 255) / 256)), dim3(256), 0, s,
                     g.NQ, (const double2*)d_part, nslab, (double2*)d_q, q_stride);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace edb
