// calq_batched.hip -- b instances sharing one collision kernel, FP32 beta and
// spectral products, blocked compensated accumulation in FP64.
//
// Operator (QCalculator.h:231-257, per instance):
//   Q_k = sum_l beta(l, [k-l]) f_l f_[k-l]
// Two exact rearrangements cut the work by ~4x:
//   * pairing: for fixed k the terms l and m = [k-l] carry the same product
//     f_l f_m, so they are summed once with beta_s = beta(l,m) + beta(m,l)
//     (the beta tables are stored pre-summed, "paired layout");
//   * Hermitian half (optional, for inputs that are spectra of real data):
//     Q_[-k] = conj(Q_k), so only k-rows (k1,k2) with k1 > 0 or k1 == 0,
//     k2 >= 0 are evaluated and mirrored on write.
//
// Mapping (CDNA4, wave64).  One lane per instance, so beta is wave-uniform:
// it is read with scalar loads and used as an SGPR operand -- every beta
// entry is fetched once per 64 instances.  A wave owns one k-row
// (k1,k2) x all k3 (or half of them for the largest N) of 64 instances and
// walks that row's schedule of l-rows (l1,l2); for each it loads the rows
// f[l1,l2,:] and f[m1,m2,:] into registers up front (in-order vector loads
// consumed progressively), and the M x M table beta_s[l3][k3] drives a fully
// unrolled product loop whose partner index m3 = [k3-l3] is resolved at
// compile time.  beta rows are prefetched one row ahead into SGPRs behind an
// explicit wait (scalar loads return out of order, so a later wait would also
// cover the prefetch).  Each l-row's FP32 partial sums are folded into FP64
// accumulators (blocked compensation).
//
// Load balance.  The kernel is persistent: nwaves waves split the flattened
// sequence of (k-row, k3-half, instance group, l-row) steps into equal
// contiguous ranges (stream-K).  Items a wave covers completely are written
// directly; the partial items at each range end go to two per-wave slots,
// which a fixup kernel sums in wave order.  Results are deterministic.
//
// Data layouts (instance-minor, Bp = batch padded to a multiple of 64):
//   fin  float2  [M^3 centred modes][Bp]
//   q    double2 [M^3 centred modes][Bp]
//   B    float   [block][l3][k3], blocks ordered by k-row then schedule.
#include <algorithm>
#include <vector>

#include "edb_internal.hpp"
#include "ops.hpp"

namespace edb {
namespace {

constexpr int cwrapc(int N, int d) {
  // centred index of the wrapped difference d, or -1 on the Nyquist residue
  int n = (N - 1) / 2;
  int r = ((d % N) + N) % N;
  return r <= n ? r + n : (r >= N - n ? r - N + n : -1);
}

// Wave-uniform tables are read through the constant address space so the
// compiler issues scalar loads and keeps beta in SGPRs.
typedef const float __attribute__((address_space(4)))* cfloat_p;
typedef const uint32_t __attribute__((address_space(4)))* cu32_p;
typedef const int __attribute__((address_space(4)))* cint_p;
typedef const long __attribute__((address_space(4)))* clong_p;

struct BatchedArgs {
  const float* B;
  const uint32_t* sched;
  const int4* rowinfo;
  const long* prefix;
  int nitems, groups;
  long total;
  int nwaves;
  const float2* fin;
  int Bp;
  double2* q;
  double2* slots;
  int* wstart;  // [nwaves] item containing each wave's first step
  int hermitian;
};

template <int N, int KPT, int H>
__device__ __forceinline__ void row_step(cfloat_p bb, const float2* __restrict__ fa,
                                         const float2* __restrict__ fcr, int Bp, float (&pr)[KPT],
                                         float (&pi)[KPT]) {
  constexpr int n = (N - 1) / 2, M = 2 * n + 1;
  constexpr int K0 = H * KPT, K1 = (H + 1) * KPT < M ? (H + 1) * KPT : M;
  constexpr int NK = K1 - K0;
  float2 c[M], a[M];
#pragma unroll
  for (int m = 0; m < M; ++m) c[m] = fcr[(long)m * Bp];
#pragma unroll
  for (int m = 0; m < M; ++m) a[m] = fa[(long)m * Bp];
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    pr[j] = 0.f;
    pi[j] = 0.f;
  }
  float bcur[NK], bnext[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) bcur[j] = bb[K0 + j];
#pragma unroll
  for (int l3 = 0; l3 < M; ++l3) {
    // Scalar loads complete out of order, so the only wait is lgkmcnt(0):
    // wait for row l3 here, then issue row l3+1 behind a scheduling barrier
    // so that the wait the compiler inserts for row l3 precedes the prefetch.
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_sched_barrier(0);
    if (l3 + 1 < M) {
#pragma unroll
      for (int j = 0; j < NK; ++j) bnext[j] = bb[(l3 + 1) * M + K0 + j];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      const int k3 = K0 + j;
      const int m = cwrapc(N, k3 - l3);
      if (m >= 0) {
        const float b = bcur[j];
        const float xr = a[l3].x * c[m].x - a[l3].y * c[m].y;
        const float xi = a[l3].x * c[m].y + a[l3].y * c[m].x;
        pr[j] += b * xr;
        pi[j] += b * xi;
      }
    }
#pragma unroll
    for (int j = 0; j < NK; ++j) bcur[j] = bnext[j];
  }
}

template <int N, int KPT>
__device__ __forceinline__ void write_row(const BatchedArgs& A, int k1c, int k2c, int h, int g, int lane,
                                          const double (&ar)[KPT], const double (&ai)[KPT]) {
  constexpr int n = (N - 1) / 2, M = 2 * n + 1;
  const bool mirror = A.hermitian && !(k1c == n && k2c == n);
  const long col = (long)g * 64 + lane;
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int k3 = h * KPT + j;
    if (k3 < M) {
      A.q[(((long)k1c * M + k2c) * M + k3) * A.Bp + col] = make_double2(ar[j], ai[j]);
      if (mirror)
        A.q[(((long)(M - 1 - k1c) * M + (M - 1 - k2c)) * M + (M - 1 - k3)) * A.Bp + col] =
            make_double2(ar[j], -ai[j]);
    }
  }
}

template <int N, int KS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void calq_batched_kernel(
    BatchedArgs A) {
  constexpr int n = (N - 1) / 2, M = 2 * n + 1, KPT = (M + KS - 1) / KS;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (w >= A.nwaves) return;
  const long a0 = A.total * w / A.nwaves, b0 = A.total * (w + 1) / A.nwaves;
  const clong_p prefix = (clong_p)A.prefix;
  const cint_p rowinfo = (cint_p)A.rowinfo;
  const cu32_p sched = (cu32_p)A.sched;
  const cfloat_p Bc = (cfloat_p)A.B;
  int lo = 0, hi = A.nitems;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= a0) lo = mid;
    else hi = mid;
  }
  int it = lo;
  if (lane == 0) A.wstart[w] = it;
  long pos = a0;
  bool first = true;
  while (pos < b0) {
    const long p0 = prefix[it], p1 = prefix[it + 1];
    const long s0 = pos - p0, s1 = (b0 < p1 ? b0 : p1) - p0;
    const int r = it / (KS * A.groups), h = (it / A.groups) % KS, g = it % A.groups;
    const int rk1 = rowinfo[4 * r], rk2 = rowinfo[4 * r + 1], rblk0 = rowinfo[4 * r + 2];
    double ar[KPT], ai[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      ar[j] = 0.;
      ai[j] = 0.;
    }
    const float2* fbase = A.fin + (long)g * 64 + lane;
    for (long s = s0; s < s1; ++s) {
      const long blk = rblk0 + s;
      const uint32_t e = sched[blk];
      const int l1c = e & 255, l2c = (e >> 8) & 255, m1c = (e >> 16) & 255, m2c = e >> 24;
      const float2* fa = fbase + (long)((l1c * M + l2c) * M) * A.Bp;
      const float2* fc = fbase + (long)((m1c * M + m2c) * M) * A.Bp;
      const cfloat_p bb = Bc + blk * (long)(M * M);
      float pr[KPT], pi[KPT];
      if (KS == 1 || h == 0) row_step<N, KPT, 0>(bb, fa, fc, A.Bp, pr, pi);
      else row_step<N, KPT, (KS > 1 ? 1 : 0)>(bb, fa, fc, A.Bp, pr, pi);
#pragma unroll
      for (int j = 0; j < KPT; ++j) {
        ar[j] += (double)pr[j];
        ai[j] += (double)pi[j];
      }
    }
    if (s0 == 0 && s1 == p1 - p0) {
      write_row<N, KPT>(A, rk1, rk2, h, g, lane, ar, ai);
    } else {
      double2* sl = A.slots + ((long)(2 * w + (first ? 0 : 1)) * KPT) * 64 + lane;
#pragma unroll
      for (int j = 0; j < KPT; ++j) sl[j * 64] = make_double2(ar[j], ai[j]);
    }
    first = false;
    pos = p0 + s1;
    ++it;
  }
}

__device__ inline long range_start(long total, int nwaves, int w) { return total * w / nwaves; }

__device__ inline int wave_of(long total, int nwaves, long x) {
  int w = (int)((x * nwaves) / total);
  if (w >= nwaves) w = nwaves - 1;
  while (w + 1 < nwaves && range_start(total, nwaves, w + 1) <= x) ++w;
  while (w > 0 && range_start(total, nwaves, w) > x) --w;
  return w;
}

// One block per interior wave boundary w (1..nwaves-1).  The item that
// contains the boundary strictly inside it (recorded by the main kernel as
// wave w's first item) was split between waves; the block of the first such
// boundary in the item sums that item's partial slots in wave order, its four
// waves splitting the k3 entries.  Items with no interior boundary were
// written by their wave.
__global__ __launch_bounds__(256) void calq_fixup_kernel(BatchedArgs A, int N, int KS) {
  const int n = (N - 1) / 2, M = 2 * n + 1, KPT = (M + KS - 1) / KS;
  const int wb = blockIdx.x + 1;
  const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  if (wb >= A.nwaves) return;
  const long p = range_start(A.total, A.nwaves, wb);
  const int it = A.wstart[wb];
  const long lo = A.prefix[it], hi = A.prefix[it + 1];
  if (lo == p || range_start(A.total, A.nwaves, wb - 1) > lo) return;
  const int wf = wave_of(A.total, A.nwaves, lo), wl = wave_of(A.total, A.nwaves, hi - 1);
  const int r = it / (KS * A.groups), h = (it / A.groups) % KS, g = it % A.groups;
  const int4 ri = A.rowinfo[r];
  const bool mirror = A.hermitian && !(ri.x == n && ri.y == n);
  const long col = (long)g * 64 + lane;
  for (int j = sub; j < KPT; j += 4) {
    const int k3 = h * KPT + j;
    if (k3 >= M) break;
    double sr = 0., si = 0.;
    for (int w = wf; w <= wl; ++w) {
      const int slot = range_start(A.total, A.nwaves, w) >= lo ? 0 : 1;
      const double2 v = A.slots[((long)(2 * w + slot) * KPT + j) * 64 + lane];
      sr += v.x;
      si += v.y;
    }
    A.q[(((long)ri.x * M + ri.y) * M + k3) * A.Bp + col] = make_double2(sr, si);
    if (mirror)
      A.q[(((long)(M - 1 - ri.x) * M + (M - 1 - ri.y)) * M + (M - 1 - k3)) * A.Bp + col] =
          make_double2(sr, -si);
  }
}

typedef void (*BatchedFn)(BatchedArgs);

// k3 rows are split in two for the largest grids to keep 2 waves per SIMD.
constexpr int ksplit_of(int N) { return N >= 27 ? 2 : 1; }

template <int N>
BatchedFn pick(int ks) {
  return ks == ksplit_of(N) ? calq_batched_kernel<N, ksplit_of(N)> : nullptr;
}

BatchedFn lookup(int N, int ks) {
  switch (N) {
#define EDB_CASE(NN) \
  case NN:           \
    return pick<NN>(ks);
    EDB_CASE(3) EDB_CASE(4) EDB_CASE(5) EDB_CASE(6) EDB_CASE(7) EDB_CASE(8) EDB_CASE(9)
    EDB_CASE(10) EDB_CASE(11) EDB_CASE(12) EDB_CASE(13) EDB_CASE(15) EDB_CASE(16) EDB_CASE(17)
    EDB_CASE(21) EDB_CASE(23) EDB_CASE(24) EDB_CASE(25) EDB_CASE(31) EDB_CASE(32) EDB_CASE(33)
#undef EDB_CASE
    default:
      return nullptr;
  }
}

}  // namespace

int batched_ksplit(int N) { return ksplit_of(N); }

// Doubles of the stream-K partial slots of nwaves waves (2 slots of KPT x 64
// complex each); the per-wave start items follow them as ints.
size_t batched_slot_doubles(int N, int nwaves) {
  const int M = 2 * ((N - 1) / 2) + 1, ks = ksplit_of(N), kpt = (M + ks - 1) / ks;
  return (size_t)nwaves * 2 * kpt * 64 * 2;
}
bool batched_supported(int N) { return lookup(N, ksplit_of(N)) != nullptr; }

// Host-side schedule: rows ordered Hermitian half first; per row the l-rows
// (l1,l2) whose partner (m1,m2) = ([k1-l1], [k2-l2]) is a mode and which is
// not lexicographically after its partner.
int build_pair_plan(int N, std::vector<int4>* rowinfo, std::vector<uint32_t>* sched, int* nrows_half) {
  Geometry gm(N);
  const int n = gm.n, M = gm.M;
  std::vector<std::pair<int, int>> half, rest;
  for (int k1c = 0; k1c < M; ++k1c)
    for (int k2c = 0; k2c < M; ++k2c) {
      int k1 = k1c - n, k2 = k2c - n;
      if (k1 > 0 || (k1 == 0 && k2 >= 0)) half.push_back({k1c, k2c});
      else rest.push_back({k1c, k2c});
    }
  *nrows_half = (int)half.size();
  rowinfo->clear();
  sched->clear();
  auto add_row = [&](int k1c, int k2c) {
    int blk0 = (int)sched->size();
    for (int l1c = 0; l1c < M; ++l1c)
      for (int l2c = 0; l2c < M; ++l2c) {
        int m1, m2;
        if (!wrap_freq(N, k1c - l1c, &m1) || !wrap_freq(N, k2c - l2c, &m2)) continue;
        int m1c = m1 + n, m2c = m2 + n;
        if (l1c > m1c || (l1c == m1c && l2c > m2c)) continue;
        sched->push_back((uint32_t)l1c | ((uint32_t)l2c << 8) | ((uint32_t)m1c << 16) |
                         ((uint32_t)m2c << 24));
      }
    rowinfo->push_back(make_int4(k1c, k2c, blk0, (int)sched->size() - blk0));
  };
  for (auto& p : half) add_row(p.first, p.second);
  for (auto& p : rest) add_row(p.first, p.second);
  return kOk;
}

int batched_prefix(const std::vector<int4>& rowinfo, int nrows, int groups, int ks,
                   std::vector<long>* prefix) {
  prefix->assign(1, 0);
  for (int r = 0; r < nrows; ++r)
    for (int h = 0; h < ks; ++h)
      for (int g = 0; g < groups; ++g) prefix->push_back(prefix->back() + rowinfo[r].w);
  return kOk;
}

int launch_calq_batched_impl(int N, const float* d_B, const uint32_t* d_sched, const int4* d_rowinfo,
                             const long* d_prefix, int nitems, long total, const float* d_fin, int Bp,
                             int groups, int hermitian, double* d_q, double* d_slots, int nwaves,
                             hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  const int ks = batched_ksplit(N);
  BatchedFn fn = lookup(N, ks);
  if (!fn) {
    set_error("calq_batched: no compiled kernel for this N");
    return kUnsupported;
  }
  EDB_REQUIRE(Bp % 64 == 0 && groups * 64 <= Bp && groups >= 1, "calq_batched: bad batch padding");
  EDB_REQUIRE(nwaves >= 1 && total >= 1, "calq_batched: empty launch");
  if (nwaves > total) nwaves = (int)total;
  BatchedArgs A;
  A.B = d_B; A.sched = d_sched; A.rowinfo = d_rowinfo; A.prefix = d_prefix;
  A.nitems = nitems; A.groups = groups; A.total = total; A.nwaves = nwaves;
  A.fin = (const float2*)d_fin; A.Bp = Bp; A.q = (double2*)d_q; A.slots = (double2*)d_slots;
  A.wstart = (int*)(d_slots + batched_slot_doubles(N, nwaves));
  A.hermitian = hermitian;
  const int wpb = 4;
  if (ev0) EDB_HIP_TRY(hipEventRecord(ev0, s));
  hipLaunchKernelGGL(fn, dim3((nwaves + wpb - 1) / wpb), dim3(64 * wpb), 0, s, A);
  EDB_HIP_TRY(hipGetLastError());
  if (ev1) EDB_HIP_TRY(hipEventRecord(ev1, s));
  if (nwaves > 1) {
    hipLaunchKernelGGL(calq_fixup_kernel, dim3(nwaves - 1), dim3(256), 0, s, A, N, ks);
    EDB_HIP_TRY(hipGetLastError());
  }
  return kOk;
}

}  // namespace edb
