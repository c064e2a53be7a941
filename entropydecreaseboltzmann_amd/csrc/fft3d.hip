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

template <int NA>
__global__ __launch_bounds__(256) void dft_pass_kernel(const double2* __restrict__ in, double2* out,
                                                       int ld, int groups, long npencil,
                                                       long stride, int inner, long outer_stride,
                                                       const double* __restrict__ tw_, double scale) {
  const cdouble_p tw = (cdouble_p)tw_;  // [2*NA]: cos, sign*sin of 2 pi r/NA
  const int lane = threadIdx.x & 63;
  const long wid = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long p = wid / groups;
  const int g = (int)(wid % groups);
  if (p >= npencil) return;
  // pencil p: index base = (p / inner) * outer_stride + (p % inner)
  const long base = (p / inner) * outer_stride + (p % inner);
  const long col = (long)g * 64 + lane;
  double2 x[NA];
#pragma unroll
  for (int j = 0; j < NA; ++j) x[j] = in[(base + j * stride) * ld + col];
#pragma unroll 1
  for (int k = 0; k < NA; ++k) {
    double sr = 0., si = 0.;
    int r = 0;  // (j*k) mod NA, wave-uniform
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const double c = tw[2 * r], s = tw[2 * r + 1];
      sr += x[j].x * c - x[j].y * s;
      si += x[j].x * s + x[j].y * c;
      r += k;
      if (r >= NA) r -= NA;
    }
    out[(base + k * stride) * ld + col] = make_double2(sr * scale, si * scale);
  }
}

__global__ void dft_pass_generic(const double2* __restrict__ in, double2* __restrict__ out, int ld,
                                 int groups, long npencil, long stride, int inner, long outer_stride,
                                 int NA, const double* __restrict__ tw, double scale) {
  const int lane = threadIdx.x & 63;
  const long wid = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long p = wid / groups;
  const int g = (int)(wid % groups);
  if (p >= npencil) return;
  const long base = (p / inner) * outer_stride + (p % inner);
  const long col = (long)g * 64 + lane;
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
                       double);

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
    h[2 * r] = cos(a);
    h[2 * r + 1] = sign * sin(a);
  }
  double* d = nullptr;
  EDB_HIP_TRY(hipMalloc(&d, sizeof(double) * 2 * NA));
  EDB_HIP_TRY(hipMemcpy(d, h.data(), sizeof(double) * 2 * NA, hipMemcpyHostToDevice));
  g_tw[key] = d;
  if (out) *out = d;
  return kOk;
}

int run_pass(int NA, const double2* in, double2* out, int ld, int groups, long npencil, long stride,
             int inner, long outer_stride, int sign, double scale, hipStream_t s) {
  const double* tw = nullptr;
  EDB_TRY(twiddles(NA, sign, &tw));
  long waves = npencil * groups;
  const int wpb = 4;
  dim3 grid((unsigned)((waves + wpb - 1) / wpb)), block(64 * wpb);
  PassFn fn = lookup_pass(NA);
  if (fn) {
    hipLaunchKernelGGL(fn, grid, block, 0, s, in, out, ld, groups, npencil, stride, inner,
                       outer_stride, tw, scale);
  } else {
    EDB_REQUIRE(in != out, "fft3d: in-place pass needs a compiled length");
    hipLaunchKernelGGL(dft_pass_generic, grid, block, 0, s, in, out, ld, groups, npencil, stride,
                       inner, outer_stride, NA, tw, scale);
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
// pass (in-place passes need the register version).
int launch_fft3d(int Nx, int Ny, int Nz, int batch, int Bp, int sign, double scale,
                 const double* d_in, double* d_out, double* d_tmp, hipStream_t s) {
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
                     ax[a].outer, sign, sc, s));
    if (dst == tmp) {
      EDB_HIP_TRY(hipMemcpyAsync(out, tmp, sizeof(double2) * (long)Nx * Ny * Nz * Bp,
                                 hipMemcpyDeviceToDevice, s));
    }
    src = out;
  }
  return kOk;
}

}  // namespace edb
