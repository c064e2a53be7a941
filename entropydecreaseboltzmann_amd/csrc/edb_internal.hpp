// edb_internal.hpp -- shared definitions of the MI355X CalQ library.
//
// Index conventions (identical to the CPU oracle and, for odd N, to the
// reference CalQ/QCalculator.h):
//   N  FFT grid points per dimension, n = (N-1)/2, M = 2n+1 Galerkin modes.
//   centred index c in [0,M) <-> frequency c-n.
//   FFTW-order index j in [0,N) <-> frequency j (j <= n) or j-N (j > n);
//   for even N the Nyquist index N/2 carries no Galerkin mode.
//   Differences k-l wrap modulo N (QCalculator.h:204-206); a wrapped value
//   that lands on the Nyquist index (even N only) is not a mode and the term
//   is absent.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace edb {

constexpr int kMaxN = 64;      // FFT grid limit of the hand-written DFT passes
constexpr int kMaxNCalq = 33;  // largest N with a compiled batched Q kernel

void set_error(const std::string& msg);
const char* last_error();

// Status codes of the C ABI (include/edb.h).
enum Status : int {
  kOk = 0,
  kInvalid = 1,
  kNoMemory = 2,
  kHipError = 3,
  kNotReady = 4,
  kUnsupported = 5,
  kBusy = 6,
};

#define EDB_HIP_TRY(expr)                                                        \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::edb::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
      return e_ == hipErrorOutOfMemory ? ::edb::kNoMemory : ::edb::kHipError;    \
    }                                                                            \
  } while (0)

#define EDB_TRY(expr)                   \
  do {                                  \
    int s_ = (expr);                    \
    if (s_ != ::edb::kOk) return s_;    \
  } while (0)

#define EDB_REQUIRE(cond, msg)          \
  do {                                  \
    if (!(cond)) {                      \
      ::edb::set_error(msg);            \
      return ::edb::kInvalid;           \
    }                                   \
  } while (0)

struct Geometry {
  int N = 0, n = 0, M = 0;
  long NQ = 0;   // M^3 Galerkin modes
  long N3 = 0;   // N^3 grid points
  explicit Geometry(int N_ = 0) : N(N_), n((N_ - 1) / 2), M(2 * ((N_ - 1) / 2) + 1) {
    NQ = (long)M * M * M;
    N3 = (long)N * N * N;
  }
};

// Frequency wrap of a difference into [-n, n]; returns false on the
// Nyquist residue of an even grid.
__host__ __device__ inline bool wrap_freq(int N, int d, int* w) {
  int n = (N - 1) / 2;
  int r = d % N;
  if (r < 0) r += N;
  if (r <= n) { *w = r; return true; }
  if (r >= N - n) { *w = r - N; return true; }
  return false;
}

__host__ __device__ inline int fftw_index(int N, int freq) { return freq >= 0 ? freq : freq + N; }

// ---------------------------------------------------------------- kernels --
// Host-side launchers; every launcher checks its shapes before launching.

// beta tables (genb.hip)
struct BetaTables {
  int N = 0, alpha = 0, filter = 0;
  double L = 0., lambda = 0., rval = 0.;
  double* d_sig = nullptr;  // M per-dimension Jackson factors (host libm, exact)
  double* d_bmm = nullptr;  // M^3 diagonal terms F(lambda|2m|,0) sigma_m^2 rval
};
int beta_tables_init(BetaTables* t, int N, int alpha, double L, int filter);
void beta_tables_free(BetaTables* t);
int launch_genb_ref64(const BetaTables& t, double* d_B, hipStream_t s);
int launch_genb_plain64(const BetaTables& t, double* d_B, int ntile, hipStream_t s);
int launch_genb_plain32(const BetaTables& t, float* d_B, int ntile, hipStream_t s);
int launch_genb_pair32(const BetaTables& t, const int* d_rowinfo, const uint32_t* d_sched,
                       int nrows, long nblocks, float* d_B, hipStream_t s);
// Compressed beta: the F table (R x Rh doubles) and the per-mode
// (sigma product, Bmm) pairs (M^3 double2) that beta is formed from.
void compressed_beta_dims(int N, long* R, long* Rh);
// Modes per wave of the on-the-fly plain kernel: kOtfK consecutive k3 of one
// k-row up to M = 27, from M = 29 on groups of kOtfK modes from one orbit of
// the cube's symmetry group (sign changes and axis permutations of k), which
// read the same rows of F.  build_orbit_groups fills whole groups with each
// orbit's modes, the last one padded with -1, and returns the group count
// (`out` holds kOtfK entries per group).
constexpr int kOtfK = 8;
inline bool otf_orbit_layout(int N) { return 2 * ((N - 1) / 2) + 1 >= 29; }
int build_orbit_groups(int N, std::vector<int>* out);
int launch_genb_compressed(const BetaTables& t, double* d_T, double* d_SB, hipStream_t s);
int launch_calq_plain_otf(int N, double rval, const double* d_T, const double* d_SB, const double* d_dup,
                          const double* d_fc, int fc_stride, const int* d_kgrp, int ngroups, double* d_part,
                          int nslab, hipStream_t s);

// plain FP64 Q (calq_plain.hip)
int launch_build_dup(int N, const double* d_fc, int fc_stride, double* d_dup, hipStream_t s);
// Kernel variants of the stored-beta stream: 0 groups of 8 terms, f_[k-l]
// gathered from the difference table with one vector load per term; 1, 2, 3
// f_[k-l] read from a per-l-row window of the table staged in LDS, beta in
// groups of 4, 2, 8 terms.  All four sum the terms in the same order with the
// same arithmetic: bitwise equal results.
constexpr int kPlainKernels = 4;
// Resident one-wave workgroups per CU of a stream variant (0 if unknown).
int plain_blocks_per_cu(int N, int kernel, bool fp32);
int launch_calq_plain32(int N, const float* d_B, const double* d_dup, const double* d_fc, int fc_stride,
                        double* d_part, int nslab, int kernel, hipStream_t s);
int launch_calq_plain64(int N, const double* d_B, const double* d_dup, const double* d_fc,
                        int fc_stride, double* d_part, int nslab, int kernel, hipStream_t s);
int launch_reduce_plain(int N, const double* d_part, int nslab, double* d_q, int q_stride,
                        hipStream_t s);

// FFT (fft3d.hip): lanes over instances, DFT passes along each axis.
int launch_fft3d(int Nx, int Ny, int Nz, int batch, int Bp, int sign, double scale,
                 const double* d_in, double* d_out, double* d_tmp, hipStream_t s,
                 const int* mask = nullptr);
int fft_prepare(int N);

}  // namespace edb
