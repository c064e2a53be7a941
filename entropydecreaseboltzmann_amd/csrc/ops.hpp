// ops.hpp -- launchers of the O(N^3) kernels (ops.hip).
#pragma once

#include <vector>

#include "edb_internal.hpp"

namespace edb {

enum CombineOp : int {
  kOpCopy = 0,
  kOpAxpy = 1,    // X + c0*Q0
  kOpRK4 = 2,     // X + c0*(Q0 + 2Q1 + 2Q2 + Q3)
  kOpSSP3b = 3,   // 0.75X + 0.25Y + c0*Q0
  kOpSSP3c = 4,   // X/3 + 2Y/3 + c0*Q0
  kOpDufort = 5,  // c0*(X + c1*Q0 + c2*(X - Y))
};

struct EntropyFixArgs {
  long N3 = 0;
  int ld = 0;
  double2* F = nullptr;      // [N^3][ld] complex grid, real part is f
  double* W = nullptr;       // [ld][N^3] work rows (Re F staged per instance)
  double* part = nullptr;    // [entropy_chunks(N3, ld)][ld] partial entropy sums
  double* H0 = nullptr;      // [ld] previous-step entropy (updated to Hfix)
  double* avg = nullptr;     // [ld] EntropyFix average
  int* avg_flag = nullptr;   // [ld]
  int* fixed = nullptr;      // [ld] 1 where mode 2 rewrote F
  double* hist_H = nullptr;  // [rec][ld] optional history
  double* hist_Hfix = nullptr;
  int* hist_fixed = nullptr;
  const double2* Fh = nullptr;   // spectral state (FFTW order) for the mean of F
  double* hist_mean = nullptr;   // [rec][ld] mean of Re F before / after the fix
  double* hist_mean_fix = nullptr;
  int rec = 0;              // history row (used when rec_ptr is null)
  const int* rec_ptr = nullptr;  // device counter of the history row
  int hist_cap = 0;
  int eligible = 0;
  const int* steps_done = nullptr;  // per instance: eligible only when steps_done[i] > 0
  int fix_mode = 0;
  double eps = 1.0e-10;
  double vol = 0.;          // (2R)^3
  double vol_over_N3 = 0.;  // (2R)^3 / N^3
};

// Number of grid chunks of the coalesced entropy pass (the partial-sum
// buffer holds chunks * ld doubles, never more than N3 * ld).
int entropy_chunks(long N3, int ld);

int launch_combine(int op, int N, int batch, int ld, double* out, const double* X, const double* Y,
                   const double* Q0, const double* Q1, const double* Q2, const double* Q3, double c0,
                   double c1, double c2, hipStream_t s);
int launch_pack(int N, int batch, int ld, bool fp32, void* out, const double* X, const double* Q0,
                double c0, hipStream_t s);
int launch_unpack(int N, int batch, int ld, double* out, const double* Q, hipStream_t s);
int launch_gather_cols(long n, int batch, int ld, const double* src, double* dst, hipStream_t s);
int launch_scatter_cols(long n, int batch, int ld, const double* src, double* dst, hipStream_t s);
int launch_real_to_complex_cols(long n, int batch, int ld, const double* src, double* dst,
                                hipStream_t s);
int launch_complex_to_real_cols(long n, int batch, int ld, const double* src, double* dst,
                                hipStream_t s);
int launch_fill(double* p, long n, double v, hipStream_t s);
int launch_select_cols(long n, int batch, int ld, const int* flag, const double* src, double* dst,
                       hipStream_t s);
int launch_entropy_fix(const EntropyFixArgs& a, int batch, hipStream_t s);
int launch_bump(int* p, hipStream_t s);
// Move column src -> dst of a [n][ld] complex array (slot compaction).
int launch_move_col(long n, int ld, int src, int dst, double* a, hipStream_t s);
// Per-slot scalars: dst <- src for the arrays given (nullable), and
// steps[i] += 1 for i < batch (increment kernel).
int launch_move_slot_scalars(int src, int dst, double* d0, double* d1, int* i0, int* i1, hipStream_t s);
int launch_incr(int* p, int n, hipStream_t s);
// For flagged slots: H0 = avg = 1, avg_flag = steps = 0 (a fresh EntropyFix
// and driver state, EntropyFix.h:20-22, BKW.h:478).
int launch_reset_slots(int ld, const int* flag, double* H0, double* avg, int* avg_flag, int* steps,
                       hipStream_t s);
// For flagged columns: F[j][col] = (rows[col*n + j], 0).
int launch_stage_real(long n, int batch, int ld, const int* flag, const double* rows, double* F,
                      hipStream_t s);
// Real part of column col of a [n][ld] complex array into out[n].
int launch_extract_col(long n, int ld, int col, const double* a, double* out, hipStream_t s);
int launch_efix(int mode, long N3, double* d_f, double eps, double aim, double* d_st, double* d_term,
                hipStream_t s);
int launch_moments(int N, int batch, int ld, double L, const double* F, double* out, hipStream_t s);
int launch_bkw_error(int N, int batch, int ld, double L, double K, const double* F, double* out,
                     hipStream_t s);

// calq_batched.hip
int batched_ksplit(int N);
// Waves per SIMD the batched kernels are built for at grid size N.
int batched_waves_per_simd(int N);
size_t batched_slot_doubles(int N, int nwaves);
size_t batched_scratch_doubles(int N, int nwaves);

// Conservative moment correction (conserve.hip): table of the 3M-2 axis
// modes (centred index and the five moment transforms) followed by the 5x5
// inverse moment matrix; the correction of Q [M^3][ld] centred, in place.
int build_moment_table(int N, double L, std::vector<double>* tab, int* E);
int launch_moment_correct(const double* d_tab, int E, int batch, int ld, double* d_q, hipStream_t s);
bool batched_supported(int N);
int build_pair_plan(int N, std::vector<int4>* rowinfo, std::vector<uint32_t>* sched, int* nrows_half);
int batched_prefix(const std::vector<int4>& rowinfo, int nrows, int groups, int ks,
                   std::vector<long>* prefix);
int quad_rows();
int quad_waves_per_block(int N);
int build_quad_plan(int N, const std::vector<int4>& rowinfo, const std::vector<uint32_t>& sched,
                    int nrows_half, std::vector<int4>* quadinfo, std::vector<int>* rowidx,
                    std::vector<int>* ublk, int* nquads_half);
int quad_prefix(const std::vector<int4>& quadinfo, int nquads, int groups, std::vector<long>* prefix);
int launch_calq_quad_impl(int N, const float* d_B, long nbeta, const uint32_t* d_sched, const int4* d_rowinfo,
                          const int4* d_quadinfo, const int* d_rowidx, const int* d_ublk, const long* d_prefix,
                          int nitems, long total, const float* d_fin, int Bp, int groups, int hermitian,
                          double* d_q, double* d_slots, int nblocks, hipStream_t s, hipEvent_t ev0 = nullptr,
                          hipEvent_t ev1 = nullptr);
int launch_calq_batched_impl(int N, const float* d_B, long nbeta, const uint32_t* d_sched,
                             const int4* d_rowinfo, const long* d_prefix, int nitems, long total,
                             const float* d_fin, int Bp, int groups, int hermitian, double* d_q,
                             double* d_slots, int nwaves, hipStream_t s, hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr);

}  // namespace edb
