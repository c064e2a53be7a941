// fft3d.hip -- batched 3D discrete Fourier transforms with FFTW's conventions
// (fft3d.h:25-48): forward X_k = sum_j x_j exp(-2 pi i jk/N), unnormalised;
// backward with exp(+...) and, for fft3d::ifft, every entry times 1/(NxNyNz).
//
// Layout is instance-minor, [grid point][ld] complex doubles, grid point
// (i*Ny + j)*Nz + k.  One lane per instance, one thread per pencil: a pass
// along an axis loads the pencil into registers (coalesced across the 64
// instances of a wave), forms each output as a length-Na dot product with
// the twiddle row exp(+-2 pi i (jk mod Na)/Na) -- the twiddles are
// wave-uniform scalar operands -- and writes back in place.  The grids here
// are tiny (N <= 64) and the O(N^4) work per instance is far below the O(N^6)
// collision operator, so a direct DFT pass in FP64 is the simplest exact form.
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "edb_internal.hpp"

namespace edb {
namespace {

typedef const double __attribute__((address_space(4)))* cdouble_p;

// Symmetric form of the length-NA DFT: with w_r = c_r + i s_r the twiddle
// table (s carries the transform's sign), inputs j and NA-j share c and have
// opposite s, so with S_j = x_j + x_{NA-j} and D_j = x_j - x_{NA-j},
//   A_k = x_0 + (-1)^k x_{NA/2} + sum_{j=1..H} c_{jk} S_j,   B_k = sum_j s_{jk} D_j,
//   X_k = A_k + i B_k,   X_{NA-k} = A_k - i B_k          (H = (NA-1)/2),
// a quarter of the multiply-adds of the direct sum.
template <int NA>
__global__ __launch_bounds__(256) void dft_pass_kernel(const double2* in, double2* out, int ld,
                                                       int groups, long npencil, long stride, int inner,
                                                       long outer_stride, const double* __restrict__ tw_,
                                                       double scale, const int* __restrict__ mask) {
  constexpr int H = (NA - 1) / 2;
  constexpr bool EVEN = (NA % 2) == 0;
  const cdouble_p tw = (cdouble_p)tw_;  // [2*NA]: cos, sign*sin of 2 pi r/NA
  const int lane = threadIdx.x & 63;
  const long wid = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long p = wid / groups;
  const int g = (int)(wid % groups);
  if (p >= npencil) return;
  // pencil p: index base = (p / inner) * outer_stride + (p % inner)
  const long base = (p / inner) * outer_stride + (p % inner);
  const long col = (long)g * 64 + lane;
  if (mask && __ballot(mask[col] != 0) == 0) return;  // no instance of this wave is wanted
  const double2* src = in + base * ld + col;
  const double2 x0 = src[0];
  double2 xh = make_double2(0., 0.);
  if (EVEN) xh = src[(long)(NA / 2) * stride * ld];
  double2 S[H > 0 ? H : 1], D[H > 0 ? H : 1];
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    const double2 a = src[(long)j * stride * ld], b = src[(long)(NA - j) * stride * ld];
    S[j - 1] = make_double2(a.x + b.x, a.y + b.y);
    D[j - 1] = make_double2(a.x - b.x, a.y - b.y);
  }
  double2* dst = out + base * ld + col;
  {
    double sr = x0.x + xh.x, si = x0.y + xh.y;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      sr += S[j].x;
      si += S[j].y;
    }
    dst[0] = make_double2(sr * scale, si * scale);
  }
#pragma unroll 1
  for (int k = 1; k <= H; ++k) {
    const double sg = (k & 1) ? -1. : 1.;
    double ar = x0.x + sg * xh.x, ai = x0.y + sg * xh.y, br = 0., bi = 0.;
    int r = 0;  // (j*k) mod NA, wave-uniform
#pragma unroll
    for (int j = 0; j < H; ++j) {
      r += k;
      if (r >= NA) r -= NA;
      const double c = tw[2 * r], sn = tw[2 * r + 1];
      ar += S[j].x * c;
      ai += S[j].y * c;
      br += D[j].x * sn;
      bi += D[j].y * sn;
    }
    dst[(long)k * stride * ld] = make_double2((ar - bi) * scale, (ai + br) * scale);
    dst[(long)(NA - k) * stride * ld] = make_double2((ar + bi) * scale, (ai - br) * scale);
  }
  if (EVEN) {  // k = NA/2: c_{j,NA/2} = (-1)^j, s = 0
    const double sg = ((NA / 2) & 1) ? -1. : 1.;
    double sr = x0.x + sg * xh.x, si = x0.y + sg * xh.y;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const double c = ((j + 1) & 1) ? -1. : 1.;
      sr += S[j].x * c;
      si += S[j].y * c;
    }
    dst[(long)(NA / 2) * stride * ld] = make_double2(sr * scale, si * scale);
  }
}

__global__ void dft_pass_generic(const double2* __restrict__ in, double2* __restrict__ out, int ld,
                                 int groups, long npencil, long stride, int inner, long outer_stride,
                                 int NA, const double* __restrict__ tw, double scale,
                                 const int* __restrict__ mask) {
  const int lane = threadIdx.x & 63;
  const long wid = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long p = wid / groups;
  const int g = (int)(wid % groups);
  if (p >= npencil) return;
  const long base = (p / inner) * outer_stride + (p % inner);
  const long col = (long)g * 64 + lane;
  if (mask && __ballot(mask[col] != 0) == 0) return;
  for (int k = 0; k < NA; ++k) {
    double sr = 0., si = 0.;
    for (int j = 0; j < NA; ++j) {
      const int r = (int)(((long)j * k) % NA);
      const double c = tw[2 * r], s = tw[2 * r + 1];
      const double2 v = in[(base + j * stride) * ld + col];
      sr += v.x * c - v.y * s;
      si += v.x * s + v.y * c;
    }
    out[(base + k * stride) * ld + col] = make_double2(sr * scale, si * scale);
  }
}

typedef void (*PassFn)(const double2*, double2*, int, int, long, long, int, long, const double*,
                       double, const int*);

PassFn lookup_pass(int NA) {
  switch (NA) {
#define EDB_P(NN) \
  case NN:        \
    return dft_pass_kernel<NN>;
    EDB_P(1) EDB_P(2) EDB_P(3) EDB_P(4) EDB_P(5) EDB_P(6) EDB_P(7) EDB_P(8) EDB_P(9) EDB_P(10)
    EDB_P(11) EDB_P(12) EDB_P(13) EDB_P(14) EDB_P(15) EDB_P(16) EDB_P(17) EDB_P(18) EDB_P(19)
    EDB_P(20) EDB_P(21) EDB_P(22) EDB_P(23) EDB_P(24) EDB_P(25) EDB_P(26) EDB_P(27) EDB_P(28)
    EDB_P(29) EDB_P(30) EDB_P(31) EDB_P(32) EDB_P(33)
#undef EDB_P
    default:
      return nullptr;
  }
}

// Twiddle tables live for the life of the process, one per (device, NA, sign).
std::mutex g_tw_mu;
std::map<std::tuple<int, int, int>, double*> g_tw;

int twiddles(int NA, int sign, const double** out) {
  int dev = 0;
  EDB_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_tuple(dev, NA, sign);
  auto itr = g_tw.find(key);
  if (itr != g_tw.end()) {
    if (out) *out = itr->second;
    return kOk;
  }
  std::vector<double> h(2 * NA);
  for (int r = 0; r < NA; ++r) {
    double a = 2. * M_PI * (double)r / (double)NA;
    double c = cos(a), sn = sin(a);
    if (4 * r == NA || 4 * r == 3 * NA) c = 0.;  // exact zeros at the quarter turns
    if (2 * r == NA) sn = 0.;
    h[2 * r] = c;
    h[2 * r + 1] = sign * sn;
  }
  double* d = nullptr;
  EDB_HIP_TRY(hipMalloc(&d, sizeof(double) * 2 * NA));
  if (hipMemcpy(d, h.data(), sizeof(double) * 2 * NA, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("fft3d: twiddle copy to the device failed");
    (void)hipFree(d);
    return kHipError;
  }
  g_tw[key] = d;
  if (out) *out = d;
  return kOk;
}

int run_pass(int NA, const double2* in, double2* out, int ld, int groups, long npencil, long stride,
             int inner, long outer_stride, int sign, double scale, const int* mask, hipStream_t s) {
  const double* tw = nullptr;
  EDB_TRY(twiddles(NA, sign, &tw));
  long waves = npencil * groups;
  const int wpb = 4;
  dim3 grid((unsigned)((waves + wpb - 1) / wpb)), block(64 * wpb);
  PassFn fn = lookup_pass(NA);
  if (fn) {
    hipLaunchKernelGGL(fn, grid, block, 0, s, in, out, ld, groups, npencil, stride, inner,
                       outer_stride, tw, scale, mask);
  } else {
    EDB_REQUIRE(in != out, "fft3d: in-place pass needs a compiled length");
    hipLaunchKernelGGL(dft_pass_generic, grid, block, 0, s, in, out, ld, groups, npencil, stride,
                       inner, outer_stride, NA, tw, scale, mask);
  }
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace

// Create the twiddle tables of an N^3 grid ahead of time (a stream capture
// must not allocate).
int fft_prepare(int N) {
  EDB_TRY(twiddles(N, -1, nullptr));
  EDB_TRY(twiddles(N, +1, nullptr));
  return kOk;
}

// sign -1: forward; +1: backward.  scale multiplies the final pass.  d_tmp
// (same size as d_out) is needed only when an axis length has no compiled
// pass (in-place passes need the register version).  With a mask ([Bp] ints),
// waves whose 64 instances all have mask 0 are skipped and those columns of
// d_out are left undefined.
int launch_fft3d(int Nx, int Ny, int Nz, int batch, int Bp, int sign, double scale,
                 const double* d_in, double* d_out, double* d_tmp, hipStream_t s, const int* mask) {
  EDB_REQUIRE(Nx >= 1 && Ny >= 1 && Nz >= 1 && Nx <= kMaxN && Ny <= kMaxN && Nz <= kMaxN,
              "fft3d: each length must be in [1, 64]");
  EDB_REQUIRE(batch >= 1 && Bp % 64 == 0 && batch <= Bp, "fft3d: bad batch / leading dimension");
  const int groups = Bp / 64;
  const double2* in = (const double2*)d_in;
  double2* out = (double2*)d_out;
  double2* tmp = (double2*)d_tmp;
  // axis z (stride 1), then y (stride Nz), then x (stride Ny*Nz)
  struct Ax { int NA; long stride; int inner; long outer; long npencil; };
  Ax ax[3] = {
      {Nz, 1, 1, (long)Nz, (long)Nx * Ny},
      {Ny, (long)Nz, Nz, (long)Ny * Nz, (long)Nx * Nz},
      {Nx, (long)Ny * Nz, Ny * Nz, (long)Nx * Ny * Nz, (long)Ny * Nz},
  };
  const double2* src = in;
  for (int a = 0; a < 3; ++a) {
    double sc = (a == 2) ? scale : 1.0;
    double2* dst = out;
    bool compiled = lookup_pass(ax[a].NA) != nullptr;
    if (!compiled && src == out) {
      EDB_REQUIRE(tmp != nullptr, "fft3d: scratch needed for this length");
      dst = tmp;
    }
    EDB_TRY(run_pass(ax[a].NA, src, dst, Bp, groups, ax[a].npencil, ax[a].stride, ax[a].inner,
                     ax[a].outer, sign, sc, mask, s));
    if (dst == tmp) {
      EDB_HIP_TRY(hipMemcpyAsync(out, tmp, sizeof(double2) * (long)Nx * Ny * Nz * Bp,
                                 hipMemcpyDeviceToDevice, s));
    }
    src = out;
  }
  return kOk;
}

}  // namespace edb
