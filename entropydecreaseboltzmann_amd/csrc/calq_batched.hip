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
// walks that row's schedule of l-rows (l1,l2); for each it loads the partner
// row c = f[m1,m2,:] into registers up front and the row a = f[l1,l2,:] a few
// l3 ahead of its use (in-order vector loads consumed progressively), and the
// M x M table beta_s[l3][k3] drives a fully unrolled product loop whose
// partner index m3 = [k3-l3] is resolved at compile time.  beta is streamed
// into SGPRs in chunks of a fixed count of values (32, two scalar loads, at
// N = 16-24; a row at N = 25 and below 16; ~1.4 rows where the k3 range is
// split), each chunk prefetched one chunk ahead behind an explicit wait
// (scalar loads return out of order, so a later wait would also cover the
// prefetch).  Each l-row's FP32 partial sums are folded into FP64
// accumulators (blocked compensation), so every work split gives the same
// FP32 partials and differs only in the FP64 merge order.
//
// The complex product a c is formed with three real products (Gauss):
//   k1 = c.x (a.x + a.y),  Re = k1 - a.y (c.x + c.y),  Im = k1 + a.x (c.y - c.x)
// with c.x + c.y and c.y - c.x formed once per partner row and a.x + a.y once
// per l3, so a pair-term costs 5 VALU instructions (1 mul, 2 FMA for the
// product, 2 FMA into the sums) instead of 6.  Each real product is within a
// few FP32 roundings of |a| |c|, the same order as the four-product form; the
// grids that must fit three waves per SIMD and cannot hold the extra partner
// values keep the four-product form.
//
// Load balance.  The kernel is persistent: nwaves waves split the flattened
// sequence of (instance group, k-row, k3-half, l-row) steps into equal
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

#include <hip/hip_ext.h>

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
  long2* rec;   // [nwaves] first and end step of that item
  float* sink;  // null: keeps the beta prefetch loads alive (never written)
  long nbeta;   // beta blocks allocated (the prefetch never reads past them)
  int hermitian;
};

// Per-grid shape of the product step (measured at N = 16, 24 and 32):
//   * gauss_of: the three-product form needs a third partner value per mode
//     (M more VGPRs); N = 17 keeps four products to stay within the 168
//     VGPRs of three waves per SIMD;
//   * awin_of: the a row is loaded this many l3-rows ahead of its use rather
//     than all up front, which frees the registers the partner row needs; ten
//     rows of lead cover the L2 latency (N = 24: two rows 1.096 ms, six 1.075,
//     ten 1.056, sixteen 1.070), fewer where the VGPR budget is tighter;
//   * bchunk_of: beta values per scalar wait.  A row per wait where a wave
//     covers all k3; where the k3 range is split (N >= 27, 16 or 17 values per
//     row) 23 values per wait give the prefetch ~1.4 rows of lead (N = 32:
//     2.01 -> 1.77 ms).  Larger chunks do not pay: their longer lead does not
//     cover a beta miss (N = 24: 28 or 32 values 1.120 / 1.123 ms against
//     1.120 with 23, same box, no spills in the step loop) -- but there the
//     chunks' scalar loads were split by the row's skipped (Nyquist) value.
//     With it loaded too, a chunk of 32 values crossing rows is two x16 loads
//     instead of the four a 23-value row needs (x16 + x4 + x2 + x1), and at
//     N = 16-24 that pays despite more SGPR spills (256 instances, same box:
//     N = 24 1.017 -> 0.989 ms, N = 23 1.145 -> 1.122, N = 17 0.221 -> 0.216,
//     N = 21 0.667 -> 0.662, N = 16 0.0940 -> 0.0935; 24 values: N = 24
//     1.019).  At N = 25 (rows of 25) and where the k3 range is split it
//     measured slower (N = 25 +2.8%, N = 32 +8%), as did 40 values anywhere.
constexpr bool gauss_of(int N) { return N != 17; }
constexpr int awin_of(int N) {
  return N <= 13 ? 4 : N <= 16 ? 6 : N == 17 ? 2 : N == 25 ? 6 : 10;
}
constexpr int bchunk_of(int N, int nk) { return N >= 27 ? 23 : (N >= 16 && N <= 24) ? 32 : nk; }
constexpr int wpe_of(int N) { return N <= 17 ? 3 : 2; }

// Touch the next beta block while the current one is computed: one lane per
// 128-byte L2 line issues a plain vector load, so the next block's scalar
// loads hit L2 instead of waiting on HBM (beta is streamed once per launch
// when there is one instance group).  The value is folded into a register
// that is stored only through a null sink pointer.
template <int N>
__device__ __forceinline__ float touch_block(const float* __restrict__ blk, int lane) {
  constexpr int n = (N - 1) / 2, M = 2 * n + 1, LINES = (M * M * 4 + 127) / 128 + 1;
  static_assert(LINES <= 64, "block larger than one wave's prefetch");
  float v = 0.f;
  if (lane < LINES) {
    const int off = lane * 32 < M * M ? lane * 32 : M * M - 1;
    v = blk[off];
  }
  return v;
}

// f rows of the rows kernel (and of the quad kernel where the k3 range is
// split) are read with buffer loads: the row's base is a wave-uniform buffer
// resource, the lane's column a fixed VGPR offset and the mode's offset an
// SGPR, so no vector instruction computes an address (46 fewer VALU per step
// at N = 24, 0.7% of the launch).  Offsets are 32-bit: the host requires
// M * Bp * 8 < 2^31 bytes per row.
typedef float v2f_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const float2* row) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)row, (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ float2 row_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const v2f_t v = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  return make_float2(v.x, v.y);
}

// fa and fcr point at the rows' first entries of this wave's instance group;
// with BUF they are wave-uniform (the lane is added here), without it they
// already include the lane.
template <int N, int KPT, int H, bool BUF>
__device__ __forceinline__ void row_step(cfloat_p bb, const float2* __restrict__ fa,
                                         const float2* __restrict__ fcr, int Bp, float (&pr)[KPT],
                                         float (&pi)[KPT], const float* __restrict__ bnext_blk, int lane,
                                         float& touched) {
  constexpr int n = (N - 1) / 2, M = 2 * n + 1;
  constexpr int K0 = H * KPT, K1 = (H + 1) * KPT < M ? (H + 1) * KPT : M;
  constexpr int NK = K1 - K0;
  constexpr bool G = gauss_of(N);
  constexpr int AW = awin_of(N) < M ? awin_of(N) : M;
  constexpr int NT = M * NK;  // beta values of this wave's part of the block
  constexpr int BC = bchunk_of(N, NK) < NT ? bchunk_of(N, NK) : NT;
  // partner row per mode: (x, x + y, y - x) in the three-product form,
  // (x, y) in cx, ce otherwise (cd unused)
  float cx[M], ce[M], cd[M];
  float2 a[M];
  const int voff = lane * 8, sstr = Bp * 8;
  // the mode's row offset is formed next to each load (one scalar multiply)
  // instead of being kept live across the step loop: the 23 offsets the a
  // and c rows share would otherwise hold SGPRs the beta chunks need (N = 24:
  // 57 -> 41 SGPR spills, none left inside the step; 0.9% faster)
  auto soff = [&](int m) {
    int r;
    asm volatile("s_mul_i32 %0, %1, %2" : "=s"(r) : "s"(sstr), "s"(m));
    return r;
  };
  auto ldc = [&](int m) { return BUF ? row_load(row_rsrc(fcr), voff, soff(m)) : fcr[(long)m * Bp]; };
  auto lda = [&](int m) { return BUF ? row_load(row_rsrc(fa), voff, soff(m)) : fa[(long)m * Bp]; };
  {
    float2 c[M];
#pragma unroll
    for (int m = 0; m < M; ++m) c[m] = ldc(m);
#pragma unroll
    for (int m = 0; m < AW; ++m) a[m] = lda(m);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      cx[m] = c[m].x;
      ce[m] = G ? c[m].x + c[m].y : c[m].y;
      cd[m] = G ? c[m].y - c[m].x : 0.f;
    }
  }
  // issued after the f rows, so their waits do not cover it; consumed at the end
  const float tv = bnext_blk ? touch_block<N>(bnext_blk, lane) : 0.f;
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    pr[j] = 0.f;
    pi[j] = 0.f;
  }
  // beta value t of this wave's part: row t / NK, column K0 + t % NK
  auto beta_at = [&](int t) { return bb[(t / NK) * M + K0 + (t % NK)]; };
  float bcur[BC], bnext[BC];
#pragma unroll
  for (int i = 0; i < BC; ++i) bcur[i] = beta_at(i);
#pragma unroll
  for (int l3 = 0; l3 < M; ++l3) {
    if (l3 + AW < M) {
      __builtin_amdgcn_sched_barrier(0);
      a[l3 + AW] = lda(l3 + AW);
      __builtin_amdgcn_sched_barrier(0);
    }
    const float sa = a[l3].x + a[l3].y;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      const int t = l3 * NK + j;
      if (t % BC == 0) {
        // Scalar loads complete out of order, so the only wait is
        // lgkmcnt(0): wait for chunk t / BC here, then issue the next chunk
        // behind a scheduling barrier so that the wait the compiler inserts
        // for this chunk precedes the prefetch.
        if (t > 0) {
#pragma unroll
          for (int i = 0; i < BC; ++i) bcur[i] = bnext[i];
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);
        if (BC % 8 == 0) {
          // At even N one value per row belongs to the skipped (Nyquist)
          // term.  Unused, it would not be loaded and would split the chunk's
          // loads; an empty asm reading it (here, after the wait, so it adds
          // no wait) keeps the chunk two whole x16 loads.
#pragma unroll
          for (int i = 0; i < BC; ++i) {
            const int ti = t + i;
            if (ti < NT && cwrapc(N, K0 + ti % NK - ti / NK) < 0) asm volatile("" ::"s"(bcur[i]));
          }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (t + BC < NT) {
#pragma unroll
          for (int i = 0; i < BC; ++i) bnext[i] = beta_at(t + BC + i < NT ? t + BC + i : NT - 1);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      const int k3 = K0 + j;
      const int m = cwrapc(N, k3 - l3);
      if (m >= 0) {
        // the roundings are written out (products and FMAs as separate
        // builtins) so that every kernel built on this step computes the same
        // FP32 partials whatever code surrounds it
        const float b = bcur[t % BC];
        float xr, xi;
        if (G) {
          const float k1 = cx[m] * sa;
          xr = __builtin_fmaf(-a[l3].y, ce[m], k1);
          xi = __builtin_fmaf(a[l3].x, cd[m], k1);
        } else {
          xr = __builtin_fmaf(a[l3].x, cx[m], -(a[l3].y * ce[m]));
          xi = __builtin_fmaf(a[l3].x, ce[m], a[l3].y * cx[m]);
        }
        pr[j] = __builtin_fmaf(b, xr, pr[j]);
        pi[j] = __builtin_fmaf(b, xi, pi[j]);
      }
    }
  }
  touched = tv;
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
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(wpe_of(N), wpe_of(N)))) void calq_batched_kernel(
    BatchedArgs A) {
  constexpr int n = (N - 1) / 2, M = 2 * n + 1, KPT = (M + KS - 1) / KS;
  const int lane = threadIdx.x & 63;
  // Workgroups are dealt to the 8 XCDs in turn; logical block (b % 8) *
  // per_xcd + b / 8 (the grid is a multiple of 8) gives each XCD a
  // contiguous run of the ranges, and with the items in instance-group-major
  // order that run stays within one or two groups' f (6 MB each at N = 24)
  // instead of all of them, so more of its row loads hit the XCD's L2.
  const int per_xcd = gridDim.x / 8;
  const int lb = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  const int w = __builtin_amdgcn_readfirstlane(lb * (blockDim.x >> 6) + (threadIdx.x >> 6));
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
  if (lane == 0) {
    A.wstart[w] = it;
    A.rec[w] = make_long2(prefix[it], prefix[it + 1]);
  }
  long pos = a0;
  bool first = true;
  float touched = 0.f;
  while (pos < b0) {
    const long p0 = prefix[it], p1 = prefix[it + 1];
    const long s0 = pos - p0, s1 = (b0 < p1 ? b0 : p1) - p0;
    const int per_g = A.nitems / A.groups;  // items of one instance group: (row, k3 part)
    const int g = it / per_g, r = (it % per_g) / KS, h = it % KS;
    const int rk1 = rowinfo[4 * r], rk2 = rowinfo[4 * r + 1], rblk0 = rowinfo[4 * r + 2];
    double ar[KPT], ai[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      ar[j] = 0.;
      ai[j] = 0.;
    }
    const float2* fbase = A.fin + (long)g * 64;  // wave-uniform: row_step adds the lane
    for (long s = s0; s < s1; ++s) {
      const long blk = rblk0 + s;
      const uint32_t e = sched[blk];
      const int l1c = e & 255, l2c = (e >> 8) & 255, m1c = (e >> 16) & 255, m2c = e >> 24;
      const float2* fa = fbase + (long)((l1c * M + l2c) * M) * A.Bp;
      const float2* fc = fbase + (long)((m1c * M + m2c) * M) * A.Bp;
      const cfloat_p bb = Bc + blk * (long)(M * M);
      float pr[KPT], pi[KPT];
      float tv = 0.f;
      const float* nb = blk + 1 < A.nbeta ? A.B + (blk + 1) * (long)(M * M) : nullptr;
      if (KS == 1 || h == 0) row_step<N, KPT, 0, true>(bb, fa, fc, A.Bp, pr, pi, nb, lane, tv);
      else row_step<N, KPT, (KS > 1 ? 1 : 0), true>(bb, fa, fc, A.Bp, pr, pi, nb, lane, tv);
      touched += tv;
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
  if (A.sink) A.sink[(long)w * 64 + lane] = touched;
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
  constexpr int kMaxCover = 64;
  __shared__ long soff[kMaxCover];  // slot base of each covering wave, in double2 units
  const int n = (N - 1) / 2, M = 2 * n + 1, KPT = (M + KS - 1) / KS;
  const int wb = blockIdx.x + 1;
  if (wb >= A.nwaves) return;
  // the item containing wave wb's first step p, as the main kernel recorded it
  const long2 rc = A.rec[wb];
  const long lo = rc.x, hi = rc.y;
  const long p = range_start(A.total, A.nwaves, wb), pprev = range_start(A.total, A.nwaves, wb - 1);
  if (lo == p || pprev > lo) return;
  const int it = A.wstart[wb];
  // wb is the first boundary inside the item, so wave wb - 1 holds its start
  // (its first slot if it starts there, else its second); the item ends in
  // the last wave whose range starts before hi
  const int wf = wb - 1;
  int wl = wb;
  while (wl + 1 < A.nwaves && range_start(A.total, A.nwaves, wl + 1) < hi) ++wl;
  const int nw = wl - wf + 1;
  for (int i = threadIdx.x; i < nw && i < kMaxCover; i += blockDim.x)
    soff[i] = (long)(2 * (wf + i) + (i == 0 && pprev < lo ? 1 : 0)) * KPT * 64;
  __syncthreads();
  const int per_g = A.nitems / A.groups;
  const int g = it / per_g, r = (it % per_g) / KS, h = it % KS;
  const int4 ri = A.rowinfo[r];
  const bool mirror = A.hermitian && !(ri.x == n && ri.y == n);
  // one thread per (k3 entry, lane), the entries spread over blockIdx.y so
  // that many blocks' loads are in flight at once; the covering waves' loads
  // of an entry are issued together
  for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < KPT * 64; e += blockDim.x * gridDim.y) {
    const int j = e >> 6, lane = e & 63;
    const int k3 = h * KPT + j;
    if (k3 >= M) continue;
    const long rel = (long)j * 64 + lane;
    double sr = 0., si = 0.;
    int i = 0;
    for (; i + 4 <= nw && i + 4 <= kMaxCover; i += 4) {
      const double2 v0 = A.slots[soff[i] + rel], v1 = A.slots[soff[i + 1] + rel];
      const double2 v2 = A.slots[soff[i + 2] + rel], v3 = A.slots[soff[i + 3] + rel];
      sr += v0.x;
      si += v0.y;
      sr += v1.x;
      si += v1.y;
      sr += v2.x;
      si += v2.y;
      sr += v3.x;
      si += v3.y;
    }
    for (; i < nw; ++i) {  // beyond the table: recompute the slot (long items only)
      long base;
      if (i < kMaxCover) {
        base = soff[i];
      } else {
        base = (long)(2 * (wf + i)) * KPT * 64;  // a wave inside the item: its first slot
      }
      const double2 v = A.slots[base + rel];
      sr += v.x;
      si += v.y;
    }
    const long col = (long)g * 64 + lane;
    A.q[(((long)ri.x * M + ri.y) * M + k3) * A.Bp + col] = make_double2(sr, si);
    if (mirror)
      A.q[(((long)(M - 1 - ri.x) * M + (M - 1 - ri.y)) * M + (M - 1 - k3)) * A.Bp + col] =
          make_double2(sr, -si);
  }
}

// ------------------------------------------------------------------ quads
// Workgroup schedule for f-row reuse.  A workgroup of Q x KS waves takes a
// quad of Q consecutive k-rows (all k3 parts): wave (j, h) computes k-row j,
// k3 part h.  All its waves walk the union of the Q rows' l-row schedules in
// the same (l1, l2) order: at a step the a-row f[l] is the same for every
// wave and the partner rows f[k_j - l] of neighbouring k-rows are neighbours
// that slide by one row per step, so the rows a workgroup touches are reused
// from L2 instead of each wave streaming its own two rows from the fabric.
// The union table holds, per step, the beta block of each k-row (-1 where
// the pair is not in that row's schedule; the wave then skips the step).
// Load balance is stream-K over (quad, instance group, union step) at
// workgroup granularity, with the partial items at range ends summed by
// calq_quad_fixup_kernel in block order.
constexpr int kQuad = 4;

struct QuadArgs {
  const float* B;
  const uint32_t* sched;
  const int4* quadinfo;  // k1c, k2c of the first row (unused), rows in the quad, union offset
  const int* rowidx;     // [quad][kQuad] rowinfo index of each row (-1 past the end)
  const int* ublk;       // [union step][kQuad] beta block or -1
  const long* prefix;
  int nitems, groups;
  long total;
  int nblocks;
  const float2* fin;
  int Bp;
  double2* q;
  double2* slots;
  int* bstart;  // [nblocks] item containing each block's first step
  float* sink;  // null: keeps the beta prefetch loads alive (never written)
  long nbeta;   // beta blocks allocated (the prefetch never reads past them)
  const int4* rowinfo;
  int hermitian;
};

template <int N, int KS>
__global__ __launch_bounds__(64 * kQuad * KS) __attribute__((amdgpu_waves_per_eu(wpe_of(N), wpe_of(N)))) void calq_quad_kernel(
    QuadArgs A) {
  constexpr int n = (N - 1) / 2, M = 2 * n + 1, KPT = (M + KS - 1) / KS;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = wave % kQuad, h = wave / kQuad;
  // XCD-contiguous logical block (the grid is a multiple of 8), as the rows kernel
  const int per_xcd = gridDim.x / 8;
  const int blk = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (blk >= A.nblocks) return;
  const long a0 = A.total * blk / A.nblocks, b0 = A.total * (blk + 1) / A.nblocks;
  const clong_p prefix = (clong_p)A.prefix;
  const cint_p quadinfo = (cint_p)A.quadinfo;
  const cint_p rowidx = (cint_p)A.rowidx;
  const cint_p rowinfo = (cint_p)A.rowinfo;
  const cint_p ublk = (cint_p)A.ublk;
  const cu32_p sched = (cu32_p)A.sched;
  const cfloat_p Bc = (cfloat_p)A.B;
  int lo = 0, hi = A.nitems;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= a0) lo = mid;
    else hi = mid;
  }
  int it = lo;
  if (threadIdx.x == 0) A.bstart[blk] = it;
  long pos = a0;
  bool first = true;
  float touched = 0.f;
  while (pos < b0) {
    const long p0 = prefix[it], p1 = prefix[it + 1];
    const long s0 = pos - p0, s1 = (b0 < p1 ? b0 : p1) - p0;
    const int per_g = A.nitems / A.groups;  // quads of one instance group
    const int g = it / per_g, qd = it % per_g;
    const int nrows = quadinfo[4 * qd + 2], uoff = quadinfo[4 * qd + 3];
    double ar[KPT], ai[KPT];
#pragma unroll
    for (int t = 0; t < KPT; ++t) {
      ar[t] = 0.;
      ai[t] = 0.;
    }
    if (j < nrows) {
      // f rows by buffer loads where the k3 range is split (N >= 27: 2-4%
      // faster at N = 32); vector addresses elsewhere (1-2% faster there)
      constexpr bool QBUF = KS > 1;
      const float2* fbase = A.fin + (long)g * 64 + (QBUF ? 0 : lane);
      for (long st = s0; st < s1; ++st) {
        const int b = ublk[(uoff + st) * kQuad + j];
        if (b < 0) continue;
        const uint32_t e = sched[b];
        const int l1c = e & 255, l2c = (e >> 8) & 255, m1c = (e >> 16) & 255, m2c = e >> 24;
        const float2* fa = fbase + (long)((l1c * M + l2c) * M) * A.Bp;
        const float2* fc = fbase + (long)((m1c * M + m2c) * M) * A.Bp;
        const cfloat_p bb = Bc + (long)b * (M * M);
        float pr[KPT], pi[KPT];
        float tv = 0.f;
        // this row's blocks are consecutive in its schedule: b + 1 is its next one
        const float* nb = b + 1 < A.nbeta ? A.B + (b + 1) * (long)(M * M) : nullptr;
        if (KS == 1 || h == 0) row_step<N, KPT, 0, QBUF>(bb, fa, fc, A.Bp, pr, pi, nb, lane, tv);
        else row_step<N, KPT, (KS > 1 ? 1 : 0), QBUF>(bb, fa, fc, A.Bp, pr, pi, nb, lane, tv);
        touched += tv;
#pragma unroll
        for (int t = 0; t < KPT; ++t) {
          ar[t] += (double)pr[t];
          ai[t] += (double)pi[t];
        }
      }
    }
    if (s0 == 0 && s1 == p1 - p0) {
      if (j < nrows) {
        const int r = rowidx[qd * kQuad + j];
        BatchedArgs W;
        W.q = A.q;
        W.Bp = A.Bp;
        W.hermitian = A.hermitian;
        write_row<N, KPT>(W, rowinfo[4 * r], rowinfo[4 * r + 1], h, g, lane, ar, ai);
      }
    } else {
      double2* sl = A.slots + ((long)(2 * blk + (first ? 0 : 1)) * (kQuad * KS) + wave) * KPT * 64 + lane;
#pragma unroll
      for (int t = 0; t < KPT; ++t) sl[t * 64] = make_double2(ar[t], ai[t]);
    }
    first = false;
    pos = p0 + s1;
    ++it;
  }
  if (A.sink) A.sink[((long)blk * (kQuad * KS) + wave) * 64 + lane] = touched;
}

// One workgroup (kQuad x KS waves) per interior block boundary; the item the
// boundary splits is summed, per wave (k-row j, k3 part h), over the slots of
// the blocks that cover it, by the block of the first such boundary.
__global__ __launch_bounds__(256 * 2) void calq_quad_fixup_kernel(QuadArgs A, int N, int KS) {
  constexpr int kMaxCover = 64;
  __shared__ long soff[kMaxCover];  // slot base of each covering block, in double2 units
  const int n = (N - 1) / 2, M = 2 * n + 1, KPT = (M + KS - 1) / KS, WPB = kQuad * KS;
  const int wb = blockIdx.x + 1;
  if (wb >= A.nblocks) return;
  const long p = range_start(A.total, A.nblocks, wb);
  const int it = A.bstart[wb];
  const long lo = A.prefix[it], hi = A.prefix[it + 1];
  if (lo == p || range_start(A.total, A.nblocks, wb - 1) > lo) return;
  const int bf = wave_of(A.total, A.nblocks, lo), bl = wave_of(A.total, A.nblocks, hi - 1);
  const int nb = bl - bf + 1;
  for (int i = threadIdx.x; i < nb && i < kMaxCover; i += blockDim.x) {
    const int b = bf + i;
    const int slot = range_start(A.total, A.nblocks, b) >= lo ? 0 : 1;
    soff[i] = (long)(2 * b + slot) * WPB * KPT * 64;
  }
  __syncthreads();
  const int per_g = A.nitems / A.groups;
  const int g = it / per_g, qd = it % per_g;
  const int nrows = A.quadinfo[qd].z;
  // one thread per (wave of the block, k3 entry, lane), spread over blockIdx.y
  for (int e = blockIdx.y * blockDim.x + threadIdx.x; e < WPB * KPT * 64; e += blockDim.x * gridDim.y) {
    const int wv = e / (KPT * 64), t = (e / 64) % KPT, lane = e & 63;
    const int j = wv % kQuad, h = wv / kQuad;
    const int k3 = h * KPT + t;
    if (j >= nrows || k3 >= M) continue;
    const long rel = (long)wv * KPT * 64 + (long)t * 64 + lane;
    double sr = 0., si = 0.;
    int i = 0;
    for (; i + 4 <= nb && i + 4 <= kMaxCover; i += 4) {
      const double2 v0 = A.slots[soff[i] + rel], v1 = A.slots[soff[i + 1] + rel];
      const double2 v2 = A.slots[soff[i + 2] + rel], v3 = A.slots[soff[i + 3] + rel];
      sr += v0.x;
      si += v0.y;
      sr += v1.x;
      si += v1.y;
      sr += v2.x;
      si += v2.y;
      sr += v3.x;
      si += v3.y;
    }
    for (; i < nb; ++i) {
      long base;
      if (i < kMaxCover) {
        base = soff[i];
      } else {
        const int b = bf + i;
        base = (long)(2 * b + (range_start(A.total, A.nblocks, b) >= lo ? 0 : 1)) * WPB * KPT * 64;
      }
      const double2 v = A.slots[base + rel];
      sr += v.x;
      si += v.y;
    }
    const int4 ri = A.rowinfo[A.rowidx[qd * kQuad + j]];
    const bool mirror = A.hermitian && !(ri.x == n && ri.y == n);
    const long col = (long)g * 64 + lane;
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

typedef void (*QuadFn)(QuadArgs);

template <int N>
QuadFn pick_quad(int ks) {
  return ks == ksplit_of(N) ? calq_quad_kernel<N, ksplit_of(N)> : nullptr;
}

QuadFn lookup_quad(int N, int ks) {
  switch (N) {
#define EDB_CASE(NN) \
  case NN:           \
    return pick_quad<NN>(ks);
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
int batched_waves_per_simd(int N) { return wpe_of(N); }
int quad_rows() { return kQuad; }

// Doubles of the stream-K partial slots of nwaves waves (2 slots of KPT x 64
// complex each); the per-wave start items follow them as ints.
size_t batched_slot_doubles(int N, int nwaves) {
  const int M = 2 * ((N - 1) / 2) + 1, ks = ksplit_of(N), kpt = (M + ks - 1) / ks;
  return (size_t)nwaves * 2 * kpt * 64 * 2;
}

// All of a launch's scratch: the slots, then per wave the (first, end) steps
// of its first item and that item's index.
size_t batched_scratch_doubles(int N, int nwaves) {
  return batched_slot_doubles(N, nwaves) + 2 * (size_t)nwaves + ((size_t)nwaves + 1) / 2;
}

// Slots of the quad schedule: 2 per wave of every block (same count as
// stream-K over nblocks * kQuad * KS waves).
int quad_waves_per_block(int N) { return kQuad * ksplit_of(N); }
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
  prefix->assign(1, 0);  // items in (instance group, row, k3 part) order
  for (int g = 0; g < groups; ++g)
    for (int r = 0; r < nrows; ++r)
      for (int h = 0; h < ks; ++h) prefix->push_back(prefix->back() + rowinfo[r].w);
  return kOk;
}

int launch_calq_batched_impl(int N, const float* d_B, long nbeta, const uint32_t* d_sched,
                             const int4* d_rowinfo, const long* d_prefix, int nitems, long total,
                             const float* d_fin, int Bp, int groups, int hermitian, double* d_q,
                             double* d_slots, int nwaves, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  const int ks = batched_ksplit(N);
  BatchedFn fn = lookup(N, ks);
  if (!fn) {
    set_error("calq_batched: no compiled kernel for this N");
    return kUnsupported;
  }
  EDB_REQUIRE(Bp % 64 == 0 && groups * 64 <= Bp && groups >= 1, "calq_batched: bad batch padding");
  EDB_REQUIRE(nwaves >= 1 && total >= 1, "calq_batched: empty launch");
  EDB_REQUIRE((long)(2 * ((N - 1) / 2) + 1) * Bp * 8 < (1L << 31), "calq_batched: batch too large for one launch");
  if (nwaves > total) nwaves = (int)total;
  BatchedArgs A;
  A.B = d_B; A.sched = d_sched; A.rowinfo = d_rowinfo; A.prefix = d_prefix;
  A.nitems = nitems; A.groups = groups; A.total = total; A.nwaves = nwaves;
  A.fin = (const float2*)d_fin; A.Bp = Bp; A.q = (double2*)d_q; A.slots = (double2*)d_slots;
  A.rec = (long2*)(d_slots + batched_slot_doubles(N, nwaves));
  A.wstart = (int*)(A.rec + nwaves);
  A.sink = nullptr;
  A.nbeta = nbeta;
  A.hermitian = hermitian;
  const int wpb = 4;
  const int nblk = ((nwaves + wpb - 1) / wpb + 7) / 8 * 8;  // a multiple of the 8 XCDs
  // timed launches take their start and end stamps from the dispatch itself,
  // so timing adds no marker between kernels
  if (ev0 && ev1) hipExtLaunchKernelGGL(fn, dim3(nblk), dim3(64 * wpb), 0, s, ev0, ev1, 0, A);
  else hipLaunchKernelGGL(fn, dim3(nblk), dim3(64 * wpb), 0, s, A);
  EDB_HIP_TRY(hipGetLastError());
  if (nwaves > 1) {
    const int kpt = (2 * ((N - 1) / 2) + 1 + ks - 1) / ks;
    hipLaunchKernelGGL(calq_fixup_kernel, dim3(nwaves - 1, (kpt * 64 + 255) / 256), dim3(256), 0, s, A, N, ks);
    EDB_HIP_TRY(hipGetLastError());
  }
  return kOk;
}

// Quads: chunks of kQuad consecutive rows of rowinfo's order (Hermitian half
// first, then the rest; a chunk never mixes the two), and for each quad the
// union of its rows' l-row schedules in (l1, l2) order with each row's beta
// block at every union step (-1 where the row has no such pair).
int build_quad_plan(int N, const std::vector<int4>& rowinfo, const std::vector<uint32_t>& sched,
                    int nrows_half, std::vector<int4>* quadinfo, std::vector<int>* rowidx,
                    std::vector<int>* ublk, int* nquads_half) {
  const int M = 2 * ((N - 1) / 2) + 1;
  quadinfo->clear();
  rowidx->clear();
  ublk->clear();
  auto add_range = [&](int r0, int r1) {
    for (int q0 = r0; q0 < r1; q0 += kQuad) {
      const int nr = std::min(kQuad, r1 - q0);
      const int uoff = (int)(ublk->size() / kQuad);
      int cur[kQuad] = {0, 0, 0, 0};
      for (int l1c = 0; l1c < M; ++l1c)
        for (int l2c = 0; l2c < M; ++l2c) {
          int b[kQuad] = {-1, -1, -1, -1};
          bool any = false;
          for (int jj = 0; jj < nr; ++jj) {
            const int4 ri = rowinfo[q0 + jj];
            if (cur[jj] < ri.w) {
              const uint32_t e = sched[ri.z + cur[jj]];
              if ((int)(e & 255) == l1c && (int)((e >> 8) & 255) == l2c) {
                b[jj] = ri.z + cur[jj];
                ++cur[jj];
                any = true;
              }
            }
          }
          if (any) ublk->insert(ublk->end(), b, b + kQuad);
        }
      for (int jj = 0; jj < nr; ++jj)
        if (cur[jj] != rowinfo[q0 + jj].w) {
          set_error("quad plan: schedule not in (l1, l2) order");
          return kInvalid;
        }
      quadinfo->push_back(make_int4(rowinfo[q0].x, rowinfo[q0].y, nr, uoff));
      for (int jj = 0; jj < kQuad; ++jj) rowidx->push_back(jj < nr ? q0 + jj : -1);
    }
    return kOk;
  };
  EDB_TRY(add_range(0, nrows_half));
  *nquads_half = (int)quadinfo->size();
  EDB_TRY(add_range(nrows_half, (int)rowinfo.size()));
  // sentinel: the union offset one past the last quad
  quadinfo->push_back(make_int4(0, 0, 0, (int)(ublk->size() / kQuad)));
  return kOk;
}

int quad_prefix(const std::vector<int4>& quadinfo, int nquads, int groups, std::vector<long>* prefix) {
  prefix->assign(1, 0);  // items in (instance group, quad) order
  for (int g = 0; g < groups; ++g)
    for (int qd = 0; qd < nquads; ++qd)
      prefix->push_back(prefix->back() + (quadinfo[qd + 1].w - quadinfo[qd].w));
  return kOk;
}

int launch_calq_quad_impl(int N, const float* d_B, long nbeta, const uint32_t* d_sched, const int4* d_rowinfo,
                          const int4* d_quadinfo, const int* d_rowidx, const int* d_ublk, const long* d_prefix,
                          int nitems, long total, const float* d_fin, int Bp, int groups, int hermitian,
                          double* d_q, double* d_slots, int nblocks, hipStream_t s, hipEvent_t ev0,
                          hipEvent_t ev1) {
  const int ks = batched_ksplit(N);
  QuadFn fn = lookup_quad(N, ks);
  if (!fn) {
    set_error("calq_quad: no compiled kernel for this N");
    return kUnsupported;
  }
  EDB_REQUIRE(Bp % 64 == 0 && groups * 64 <= Bp && groups >= 1, "calq_quad: bad batch padding");
  EDB_REQUIRE(nblocks >= 1 && total >= 1, "calq_quad: empty launch");
  EDB_REQUIRE((long)(2 * ((N - 1) / 2) + 1) * Bp * 8 < (1L << 31), "calq_quad: batch too large for one launch");
  if (nblocks > total) nblocks = (int)total;
  const int wpb = kQuad * ks;
  QuadArgs A;
  A.B = d_B; A.sched = d_sched; A.quadinfo = d_quadinfo; A.rowidx = d_rowidx; A.ublk = d_ublk;
  A.prefix = d_prefix; A.nitems = nitems; A.groups = groups; A.total = total; A.nblocks = nblocks;
  A.fin = (const float2*)d_fin; A.Bp = Bp; A.q = (double2*)d_q; A.slots = (double2*)d_slots;
  A.bstart = (int*)(d_slots + batched_slot_doubles(N, nblocks * wpb));
  A.rowinfo = d_rowinfo; A.hermitian = hermitian;
  A.sink = nullptr;
  A.nbeta = nbeta;
  const dim3 grid((nblocks + 7) / 8 * 8);  // a multiple of the 8 XCDs
  if (ev0 && ev1) hipExtLaunchKernelGGL(fn, grid, dim3(64 * wpb), 0, s, ev0, ev1, 0, A);
  else hipLaunchKernelGGL(fn, grid, dim3(64 * wpb), 0, s, A);
  EDB_HIP_TRY(hipGetLastError());
  if (nblocks > 1) {
    const int kpt = (2 * ((N - 1) / 2) + 1 + ks - 1) / ks;
    hipLaunchKernelGGL(calq_quad_fixup_kernel, dim3(nblocks - 1, (wpb * kpt * 64 + 64 * wpb - 1) / (64 * wpb)),
                       dim3(64 * wpb), 0, s, A, N, ks);
    EDB_HIP_TRY(hipGetLastError());
  }
  return kOk;
}

}  // namespace edb
