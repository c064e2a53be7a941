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
#include <algorithm>
#include <cstdlib>
#include <string>
#include <type_traits>

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

// The same stream with the difference table read through LDS.  For one
// l-row (l1, l2) the wave's lanes (64 consecutive centred k, spanning at most
// rq = ceil(64/M) + 1 k-rows) read D only on rows r_q = (k1_q - l1 + M-1) W +
// (k2_q - l2 + M-1) of their k-rows q, at position k3 - l3 + M-1: rq x W
// entries (2.9 KB at N = 24) instead of one 16-byte vector gather per term.
// The window of the next l-row is loaded into registers while the current
// one is summed and stored to the other LDS buffer at the row's end, so the
// vector-memory unit carries only the beta stream and a few window loads per
// row, and each term reads its f_[k-l] with one ds_read_b128.  Terms are
// summed in the same l order with the same arithmetic as
// calq_plain64_kernel: the results are bitwise identical.
constexpr int kWinMaxLoads = 6;  // window loads per lane and row: rq W <= 128 + 4M <= 6 * 64 for M <= 63

__host__ __device__ inline int win_rows(int M) { return (64 + M - 1) / M + 1; }
// Window loads per lane and row, rounded up to a compiled count (3, 4 or 6):
// the staging registers are the kernel's largest VGPR item after the beta group.
inline int win_loads(int M) {
  const int n = (win_rows(M) * (2 * M - 1) + 63) / 64;
  return n <= 3 ? 3 : (n <= 4 ? 4 : 6);
}

template <int U, int NST, typename T>
__global__ __launch_bounds__(64) void calq_plain64_lds_kernel(int M, long NQ, const T* __restrict__ B,
                                                              const double2* __restrict__ dup,
                                                              const double2* __restrict__ fc, int stride,
                                                              double2* __restrict__ part, int nslab) {
  extern __shared__ double2 win[];  // [2][rq * W]
  const int W = 2 * M - 1, rq = win_rows(M), wsz = rq * W;
  const long MM = (long)M * M;
  const int lane = threadIdx.x;
  const int tile = blockIdx.x, slab = blockIdx.y;
  const long kc = (long)tile * 64 + lane;
  const long kcc = kc < NQ ? kc : NQ - 1;  // padded lanes read a valid address; their beta is 0
  const int k3 = (int)(kcc % M);
  const long kr0 = ((long)tile * 64) / M;  // first k-row (k1 M + k2) of the wave
  const int q = (int)(kcc / M - kr0);       // this lane's k-row among the wave's, < rq
  const long l0 = NQ * slab / nslab, l1e = NQ * (slab + 1) / nslab;
  const long kp = (long)gridDim.x * 64;  // padded row length
  const T* bt = B + (long)tile * 64 + lane;
  double qr = 0., qi = 0.;
  double2 st[NST];
  auto window_load = [&](long lr) {  // D-rows of l-row lr (l1 = lr / M, l2 = lr % M) into registers
    const int l1 = (int)(lr / M), l2 = (int)(lr % M);
#pragma unroll
    for (int i = 0; i < NST; ++i) {
      const int e = lane + 64 * i;
      const int qq = e / W, p = e - qq * W;
      const long kr = kr0 + qq;
      st[i] = make_double2(0., 0.);
      if (e < wsz && kr < MM) {
        const int k1q = (int)(kr / M), k2q = (int)(kr - (long)k1q * M);
        const long drow = (long)(k1q - l1 + M - 1) * W + (k2q - l2 + M - 1);
        st[i] = dup[drow * W + p];
      }
    }
  };
  auto window_store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NST; ++i) {
      const int e = lane + 64 * i;
      if (e < wsz) win[buf * wsz + e] = st[i];
    }
  };
  long lr = l0 / M;
  const long lr_end = (l1e - 1) / M;
  window_load(lr);
  window_store(0);
  __syncthreads();
  int buf = 0;
  for (; lr <= lr_end; ++lr) {
    if (lr < lr_end) window_load(lr + 1);
    const long la = lr * M > l0 ? lr * M : l0;
    const long lb = (lr + 1) * M < l1e ? (lr + 1) * M : l1e;
    const double2* wrow = win + buf * wsz + q * W + k3 + M - 1 - (la - lr * M);  // entry of l = la
    long l = la;
    for (; l + U <= lb; l += U) {
      double b[U];
      double2 c[U], a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        b[u] = (double)__builtin_nontemporal_load(bt + (l + u) * kp);
        a[u] = fc[(l + u) * stride];
        c[u] = wrow[-(int)(l + u - la)];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) accum_term(a[u], c[u], b[u], qr, qi);
    }
    for (; l < lb; ++l) {
      const double b = (double)__builtin_nontemporal_load(bt + l * kp);
      const double2 a = fc[l * stride];
      const double2 c = wrow[-(int)(l - la)];
      accum_term(a, c, b, qr, qi);
    }
    if (lr < lr_end) {  // the next row's window into the other buffer
      window_store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
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
// Bmm[m], in genb's order, so every beta equals the stored one bit for bit.
// Nothing of size M^6 exists.  Lanes run over 64 consecutive l, so the f_l,
// f_[k-l] (difference table) and (sigma, Bmm)[m] reads are contiguous across
// lanes (m = k - l runs backwards with l); each lane accumulates its l terms
// for the wave's K modes in l order, the 64 lanes are then summed with a fixed
// butterfly and slabs of l are summed afterwards in a fixed order, so the
// result is deterministic (it differs from the stored kernel only in
// summation order).  Two ways of choosing a wave's K modes:
//
// k3 block (grids up to M = 27): K consecutive k3 of one k-row (k1, k2).  The
// K modes' f_[k-l] and (sigma, Bmm) reads of a lane fall on the same cache
// lines, and T (12 MB at N = 25) stays in L2 whichever rows are in use.
//
// Orbit (M >= 29).  For a mode k, l + m = k - N w where w is the wrap of
// k - l, so the row |l+m|^2 depends only on which axes wrap and on |k1|,
// |k2|, |k3|: the modes of one orbit of the cube's symmetries (sign changes
// and axis permutations) read the same <= 8 rows of T.  A wave takes K modes
// of one orbit (the groups come in orbit order) and a slab of l, and block b
// runs item (b mod 8) * per_xcd + b / 8: workgroups are dealt to the eight
// XCDs in turn, so each XCD walks its own contiguous run of orbits and the
// rows its waves share stay in its L2 (at N = 64 a row is 46 KB and T
// 536 MB).  The price is K separate lines per lane for the f_[k-l] and
// (sigma, Bmm) reads, which costs more than it saves while T fits in L2.
template <int K>
__global__ __launch_bounds__(64) void calq_plain_otf_k3_kernel(int N, int M, long NQ, double rval,
                                                            const double* __restrict__ T, int Rh,
                                                            const double2* __restrict__ SB,
                                                            const double2* __restrict__ dup,
                                                            const double2* __restrict__ fc, int stride,
                                                            double2* __restrict__ part, int nslab) {
  const int n = (N - 1) / 2, W = 2 * M - 1;
  const int lane = threadIdx.x;
  const int nk3 = (M + K - 1) / K;
  const int kb = blockIdx.x, slab = blockIdx.y;
  const int k1 = kb / (M * nk3), k2 = (kb / nk3) % M, k3b = (kb % nk3) * K;
  const int dk12 = ((k1 + M - 1) * W + (k2 + M - 1)) * W + (M - 1);
  const int l0 = (int)(NQ * slab / nslab), l1e = (int)(NQ * (slab + 1) / nslab);
  double ar[K], ai[K];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    ar[u] = 0.;
    ai[u] = 0.;
  }
  for (int lb = l0; lb < l1e; lb += 64) {
    const bool active = lb + lane < l1e;
    const int l = active ? lb + lane : l1e - 1;  // inactive lanes read valid addresses, add 0
    const int row = l / M, a3 = l - row * M;
    const int a1 = row / M, a2 = row - a1 * M;
    const double2 fl = fc[(long)l * stride];
    const double sl = SB[l].x;
    bool v1, v2;
    const int m1 = wrap_small(N, n, k1 - a1, &v1), m2 = wrap_small(N, n, k2 - a2, &v2);
    const int l1 = a1 - n, l2 = a2 - n, l3 = a3 - n;
    const bool ok12 = active && v1 && v2;
    const int A12 = (l1 + m1) * (l1 + m1) + (l2 + m2) * (l2 + m2);
    const int B12 = (l1 - m1) * (l1 - m1) + (l2 - m2) * (l2 - m2);
    const int mc12 = ((m1 + n) * M + (m2 + n)) * M + n;
    const int dl = dk12 - (a1 * W + a2) * W - a3;  // f_[k-l] sits at dl + k3
#pragma unroll
    for (int u = 0; u < K; ++u) {
#pragma clang fp contract(off)
      // past the row's last k3 the mode is clamped for addressing and adds 0
      const bool kin = k3b + u < M;
      const int k3 = kin ? k3b + u : M - 1;
      bool v3;
      const int m3 = wrap_small(N, n, k3 - a3, &v3);
      const int A = A12 + (l3 + m3) * (l3 + m3);
      const int Bq = B12 + (l3 - m3) * (l3 - m3);
      const double F = T[A * Rh + (Bq >> 1)];
      const double2 sb = SB[mc12 + m3];
      const double2 c = dup[dl + k3];
      const double v = F * sl * sb.x * rval - sb.y;
      const double b = (kin && ok12 && v3 && A != 0) ? v : 0.;
      accum_term(fl, c, b, ar[u], ai[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < K; ++u) {
    double r = ar[u], i = ai[u];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      r += __shfl_xor(r, o);
      i += __shfl_xor(i, o);
    }
    const int k3 = k3b + u;
    if (lane == 0 && k3 < M) part[(long)slab * NQ + ((long)k1 * M + k2) * M + k3] = make_double2(r, i);
  }
}

template <int K>
__global__ __launch_bounds__(64) void calq_plain_otf_orbit_kernel(int N, int M, long NQ, double rval,
                                                            const double* __restrict__ T, int Rh,
                                                            const double2* __restrict__ SB,
                                                            const double2* __restrict__ dup,
                                                            const double2* __restrict__ fc, int stride,
                                                            const int* __restrict__ kgrp, long nitems,
                                                            long per_xcd, int nslab,
                                                            double2* __restrict__ part) {
  const int n = (N - 1) / 2, W = 2 * M - 1;
  const int lane = threadIdx.x;
  const long b = blockIdx.x;
  const long item = (b & 7) * per_xcd + (b >> 3);
  if (item >= nitems) return;
  const long grp = item / nslab;
  const int slab = (int)(item - grp * nslab);
  int kc[K], k1[K], k2[K], k3[K], dk[K];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    const int c = __builtin_amdgcn_readfirstlane(kgrp[grp * K + u]);
    kc[u] = c;
    const int cc = c >= 0 ? c : 0;  // padding entries address mode 0 and add nothing
    k1[u] = cc / (M * M);
    k2[u] = (cc / M) % M;
    k3[u] = cc % M;
    dk[u] = ((k1[u] + M - 1) * W + (k2[u] + M - 1)) * W + (k3[u] + M - 1);
  }
  const int l0 = (int)(NQ * slab / nslab), l1e = (int)(NQ * (slab + 1) / nslab);
  double ar[K], ai[K];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    ar[u] = 0.;
    ai[u] = 0.;
  }
  for (int lb = l0; lb < l1e; lb += 64) {
    const bool active = lb + lane < l1e;
    const int l = active ? lb + lane : l1e - 1;  // inactive lanes read valid addresses, add 0
    const int row = l / M, a3 = l - row * M;
    const int a1 = row / M, a2 = row - a1 * M;
    const double2 fl = fc[(long)l * stride];
    const double sl = SB[l].x;
    const int l1 = a1 - n, l2 = a2 - n, l3 = a3 - n;
    const int ld = (a1 * W + a2) * W + a3;  // f_[k-l] sits at dk - ld
#pragma unroll
    for (int u = 0; u < K; ++u) {
#pragma clang fp contract(off)
      bool v1, v2, v3;
      const int m1 = wrap_small(N, n, k1[u] - a1, &v1), m2 = wrap_small(N, n, k2[u] - a2, &v2),
                m3 = wrap_small(N, n, k3[u] - a3, &v3);
      const int A = (l1 + m1) * (l1 + m1) + (l2 + m2) * (l2 + m2) + (l3 + m3) * (l3 + m3);
      const int Bq = (l1 - m1) * (l1 - m1) + (l2 - m2) * (l2 - m2) + (l3 - m3) * (l3 - m3);
      const double F = T[A * Rh + (Bq >> 1)];
      const double2 sb = SB[((m1 + n) * M + (m2 + n)) * M + (m3 + n)];
      const double2 c = dup[dk[u] - ld];
      const double v = F * sl * sb.x * rval - sb.y;
      const double bt = (kc[u] >= 0 && active && v1 && v2 && v3 && A != 0) ? v : 0.;
      accum_term(fl, c, bt, ar[u], ai[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < K; ++u) {
    double r = ar[u], i = ai[u];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      r += __shfl_xor(r, o);
      i += __shfl_xor(i, o);
    }
    if (lane == 0 && kc[u] >= 0) part[(long)slab * NQ + kc[u]] = make_double2(r, i);
  }
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

template <typename T>
int launch_plain_stream(const char* what, int N, const T* d_B, const double* d_dup, const double* d_fc,
                        int fc_stride, double* d_part, int nslab, int kernel, hipStream_t s) {
  Geometry g(N);
  EDB_REQUIRE(nslab >= 1 && nslab <= 4096 && nslab <= g.NQ, std::string(what) + ": bad slab count");
  EDB_REQUIRE(kernel >= 0 && kernel < kPlainKernels, std::string(what) + ": unknown kernel variant");
  const int ntile = (int)((g.NQ + 63) / 64);
  const dim3 grid(ntile, nslab);
  const size_t lds = kernel == 0 ? 0 : sizeof(double2) * 2 * win_rows(g.M) * (2 * g.M - 1);
  EDB_REQUIRE(win_rows(g.M) * (2 * g.M - 1) <= kWinMaxLoads * 64, std::string(what) + ": window too large");
  auto args = [&](auto fn) {
    hipLaunchKernelGGL(fn, grid, dim3(64), lds, s, g.M, g.NQ, d_B, (const double2*)d_dup, (const double2*)d_fc,
                       fc_stride, (double2*)d_part, nslab);
  };
  const int nst = win_loads(g.M);
  auto lds_kernel = [&](auto u) {
    constexpr int U = decltype(u)::value;
    if (nst == 3) args(calq_plain64_lds_kernel<U, 3, T>);
    else if (nst == 4) args(calq_plain64_lds_kernel<U, 4, T>);
    else args(calq_plain64_lds_kernel<U, kWinMaxLoads, T>);
  };
  switch (kernel) {
    case 0: args(calq_plain64_kernel<8, T>); break;
    case 1: lds_kernel(std::integral_constant<int, 4>()); break;
    case 2: lds_kernel(std::integral_constant<int, 2>()); break;
    default: lds_kernel(std::integral_constant<int, 8>()); break;
  }
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

// Resident workgroups per CU of the stream kernel `kernel` (one wave each):
// slab counts up to this many per CU run in a single round.
template <typename T>
int plain_occupancy(int N, int kernel) {
  Geometry g(N);
  const size_t lds = kernel == 0 ? 0 : sizeof(double2) * 2 * win_rows(g.M) * (2 * g.M - 1);
  const int nst = win_loads(g.M);
  int n = 0;
  auto occ = [&](auto fn) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, lds) != hipSuccess) n = 0;
  };
  auto lds_kernel = [&](auto u) {
    constexpr int U = decltype(u)::value;
    if (nst == 3) occ(calq_plain64_lds_kernel<U, 3, T>);
    else if (nst == 4) occ(calq_plain64_lds_kernel<U, 4, T>);
    else occ(calq_plain64_lds_kernel<U, kWinMaxLoads, T>);
  };
  switch (kernel) {
    case 0: occ(calq_plain64_kernel<8, T>); break;
    case 1: lds_kernel(std::integral_constant<int, 4>()); break;
    case 2: lds_kernel(std::integral_constant<int, 2>()); break;
    default: lds_kernel(std::integral_constant<int, 8>()); break;
  }
  (void)hipGetLastError();
  return n;
}

int plain_blocks_per_cu(int N, int kernel, bool fp32) {
  return fp32 ? plain_occupancy<float>(N, kernel) : plain_occupancy<double>(N, kernel);
}

int launch_calq_plain64(int N, const double* d_B, const double* d_dup, const double* d_fc,
                        int fc_stride, double* d_part, int nslab, int kernel, hipStream_t s) {
  return launch_plain_stream("calq_plain64", N, d_B, d_dup, d_fc, fc_stride, d_part, nslab, kernel, s);
}

// Same evaluation with the FP32 copy of beta (products and sums stay FP64):
// half the bytes of the HBM stream.
int launch_calq_plain32(int N, const float* d_B, const double* d_dup, const double* d_fc, int fc_stride,
                        double* d_part, int nslab, int kernel, hipStream_t s) {
  return launch_plain_stream("calq_plain32", N, d_B, d_dup, d_fc, fc_stride, d_part, nslab, kernel, s);
}

int build_orbit_groups(int N, std::vector<int>* out) {
  Geometry g(N);
  struct Mode {
    long key;  // |k|^2, then |k1|, |k2|, |k3| in decreasing order: one key per orbit
    int kc;    // (each |k_i| <= n <= 31 takes 6 bits: handles exist for N <= 64 only)
  };
  if (g.n >= 64) return 0;
  std::vector<Mode> modes;
  modes.reserve(g.NQ);
  for (int c = 0; c < g.NQ; ++c) {
    int a[3] = {std::abs(c / (g.M * g.M) - g.n), std::abs((c / g.M) % g.M - g.n), std::abs(c % g.M - g.n)};
    std::sort(a, a + 3);
    const long r2 = (long)a[0] * a[0] + (long)a[1] * a[1] + (long)a[2] * a[2];
    modes.push_back({((r2 * 64 + a[2]) * 64 + a[1]) * 64 + a[0], c});
  }
  std::sort(modes.begin(), modes.end(), [](const Mode& x, const Mode& y) {
    return x.key != y.key ? x.key < y.key : x.kc < y.kc;
  });
  out->clear();
  for (size_t i = 0; i < modes.size();) {
    size_t j = i;
    while (j < modes.size() && modes[j].key == modes[i].key) out->push_back(modes[j++].kc);
    while (out->size() % kOtfK) out->push_back(-1);
    i = j;
  }
  return (int)(out->size() / kOtfK);
}

int launch_calq_plain_otf(int N, double rval, const double* d_T, const double* d_SB, const double* d_dup,
                          const double* d_fc, int fc_stride, const int* d_kgrp, int ngroups, double* d_part,
                          int nslab, hipStream_t s) {
  Geometry g(N);
  EDB_REQUIRE(nslab >= 1 && nslab <= 4096 && nslab <= g.NQ, "calq_plain_otf: bad slab count");
  long R, Rh;
  compressed_beta_dims(N, &R, &Rh);
  if (!otf_orbit_layout(N)) {
    const int nkb = g.M * g.M * ((g.M + kOtfK - 1) / kOtfK);
    hipLaunchKernelGGL((calq_plain_otf_k3_kernel<kOtfK>), dim3(nkb, nslab), dim3(64), 0, s, N, g.M, g.NQ, rval,
                       d_T, (int)Rh, (const double2*)d_SB, (const double2*)d_dup, (const double2*)d_fc, fc_stride,
                       (double2*)d_part, nslab);
  } else {
    EDB_REQUIRE(d_kgrp && ngroups >= 1, "calq_plain_otf: no mode groups");
    const long nitems = (long)ngroups * nslab, per_xcd = (nitems + 7) / 8;
    EDB_REQUIRE(8 * per_xcd <= 0x7fffffffL, "calq_plain_otf: grid too large");
    hipLaunchKernelGGL((calq_plain_otf_orbit_kernel<kOtfK>), dim3((unsigned)(8 * per_xcd)), dim3(64), 0, s, N,
                       g.M, g.NQ, rval, d_T, (int)Rh, (const double2*)d_SB, (const double2*)d_dup,
                       (const double2*)d_fc, fc_stride, d_kgrp, nitems, per_xcd, nslab, (double2*)d_part);
  }
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
