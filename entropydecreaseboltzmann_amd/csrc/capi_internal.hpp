// capi_internal.hpp -- state of a QCalculator handle and the host helpers
// shared by the C ABI (capi.hip) and the request server (server.hip).
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "edb_internal.hpp"
#include "ops.hpp"

namespace edb {

template <typename T>
int dmalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return kOk;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error(std::string("hipMalloc(") + std::to_string(sizeof(T) * count) + " bytes): " +
              hipGetErrorString(e));
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? kNoMemory : kHipError;
  }
  return kOk;
}

template <typename T>
void dfree(T** p) {
  if (*p) (void)hipFree((void*)*p);
  *p = nullptr;
}

inline int pad64(int b) { return ((b + 63) / 64) * 64; }
int cu_count();

struct PrefixEntry {
  long* d_prefix = nullptr;
  int nitems = 0;
  long total = 0;
};

// Per-evaluation scratch of the Q kernels: the batched kernel's stream-K
// partial slots (and per-wave start items), the plain kernel's difference
// table and l-slab partial sums.  Kernels write it on every evaluation, so
// two streams that evaluate Q at the same time need a workspace each: a
// QCalculator handle keeps one for its own calls and every solver owns one,
// also when it shares the handle's collision kernel.  Buffers that grow are
// retired, not freed, because a HIP graph captured earlier may still name
// them; ws_free releases everything once no graph or launch uses them.
struct Workspace {
  double* d_slots = nullptr;
  size_t slots_count = 0;  // doubles
  double* d_dup = nullptr;
  double* d_part = nullptr;
  long part_slabs = 0;  // capacity of d_part, in slabs of M^3 complex
  std::vector<void*> retired;
};
void ws_free(Workspace* w);

}  // namespace edb

struct edb_qcalc {
  int N = 0, alpha = 0, filter = 1;
  double L = 0.;
  edb::Geometry g;
  hipStream_t stream = nullptr;
  edb::BetaTables tables;
  bool tables_ready = false;
  // plain FP64 (and its FP32 copy)
  double* d_plain = nullptr;
  float* d_plain32 = nullptr;
  // compressed beta: F table over (|l+m|^2, |l-m|^2 / 2) and (sigma, Bmm) per mode
  double* d_ftab = nullptr;
  double* d_sb = nullptr;
  int* d_kgrp = nullptr;  // modes in groups of kOtfK from one symmetry orbit (-1 pads)
  int n_kgrp = 0;         // number of groups
  int ntile = 0;
  int nslab = 16;    // l-slabs of the FP64 plain kernel
  int nslab32 = 0;   // l-slabs of the FP32-beta plain kernel (0: nslab)
  int pkern = 1;     // stored-beta stream variant of the FP64 plain kernel (launch_calq_plain64)
  int pkern32 = 1;   // same for the FP32-beta plain kernel
  // scratch of the handle's own evaluations (calQ*, time_*, autotune)
  edb::Workspace ws;
  // paired FP32
  float* d_pair = nullptr;
  int4* d_rowinfo = nullptr;
  uint32_t* d_sched = nullptr;
  std::vector<int4> rowinfo;
  int nrows_half = 0;
  long nblocks = 0;
  std::map<std::pair<int, int>, edb::PrefixEntry> prefixes;  // (groups, hermitian + 2 * schedule)
  std::mutex prefix_mu;
  int nwaves = 0;
  // quad schedule (schedule 1): workgroups of 4 neighbouring k-rows
  int schedule = 0;
  int nblocks_quad = 0;
  std::vector<int4> quadinfo;
  int nquads_half = 0;
  int4* d_quadinfo = nullptr;
  int* d_rowidx = nullptr;
  int* d_ublk = nullptr;
  // conservative moment correction
  int conserve = 0;
  double* d_mom = nullptr;
  int mom_E = 0;
  std::mutex mu;
};

namespace edb {

int ensure_tables(edb_qcalc* q);
int get_prefix(edb_qcalc* q, int groups, int hermitian, PrefixEntry** out);
int ensure_slots(Workspace* ws, int N, int nwaves);
// Persistent waves (rows) or workgroups (quads) of the batched kernel for
// `groups` instance groups under the handle's schedule, before any cap by the
// schedule's item count.
int batched_count(const edb_qcalc* q);
// Launch shape of the batched kernel for `groups` instance groups under the
// handle's schedule: its prefix and persistent waves (rows) or workgroups
// (quads), with the partial-sum slots allocated in ws.
int batched_shape(edb_qcalc* q, Workspace* ws, int groups, int hermitian, PrefixEntry** pe, int* count);
// Device table of the moment correction for q's (N, L), built on first use.
int ensure_moment_table(edb_qcalc* q);
// Batched Q on device arrays already in the kernel layout (float2 [M^3][ld] in,
// double2 [M^3][ld] out); events, if given, bracket the main kernel.
int batched_device(edb_qcalc* q, Workspace* ws, int batch, int ld, int hermitian, double* d_q,
                   const float* d_fin, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// Plain-path scratch: the difference table and room for `slabs` partial sums.
int ensure_plain_scratch(const edb_qcalc* q, Workspace* ws, int slabs);
// l-slabs of the plain kernel for a PlainSource.
int plain_slabs(const edb_qcalc* q, int src);
// Where the plain kernel takes beta from, and the plain Q of one column
// (centred FP64 input with stride fc_stride) with that beta.
enum PlainSource { kPlainFP64 = 0, kPlainFP32 = 1, kPlainCompressed = 2 };
int plain_column(edb_qcalc* q, Workspace* ws, const double* d_fc, int fc_stride, double* d_q, int q_stride,
                 hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, int src = kPlainFP64);

}  // namespace edb
