// capi.hip -- the C ABI (include/edb.h): QCalculator, fft3d, EntropyFix and
// the BKW3D drivers, as host C++ over the HIP kernels of this directory.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/edb.h"
#include "capi_internal.hpp"
#include "edb_internal.hpp"
#include "ops.hpp"

namespace edb {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

int cu_count() {
  static int cached = 0;
  if (cached) return cached;
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    n = 256;
  cached = n;
  return n;
}

}  // namespace edb

using namespace edb;


namespace edb {

int ensure_tables(edb_qcalc* q) {
  if (q->tables_ready) return kOk;
  EDB_TRY(beta_tables_init(&q->tables, q->N, q->alpha, q->L, q->filter));
  q->tables_ready = true;
  return kOk;
}

int get_prefix(edb_qcalc* q, int groups, int hermitian, PrefixEntry** out) {
  // solvers sharing q may ask from several threads; entries are never removed
  // before destroy, so the pointer stays valid after the lock is released
  std::lock_guard<std::mutex> lk(q->prefix_mu);
  auto key = std::make_pair(groups, hermitian + 2 * q->schedule);
  auto it = q->prefixes.find(key);
  if (it == q->prefixes.end()) {
    std::vector<long> pre;
    if (q->schedule == 1) {
      int nquads = hermitian ? q->nquads_half : (int)q->quadinfo.size() - 1;
      quad_prefix(q->quadinfo, nquads, groups, &pre);
    } else {
      int nrows = hermitian ? q->nrows_half : (int)q->rowinfo.size();
      batched_prefix(q->rowinfo, nrows, groups, batched_ksplit(q->N), &pre);
    }
    PrefixEntry e;
    e.nitems = (int)pre.size() - 1;
    e.total = pre.back();
    EDB_TRY(dmalloc(&e.d_prefix, pre.size()));
    if (hipMemcpy(e.d_prefix, pre.data(), sizeof(long) * pre.size(), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("batched prefix: copy to the device failed");
      dfree(&e.d_prefix);
      return kHipError;
    }
    it = q->prefixes.emplace(key, e).first;
  }
  *out = &it->second;
  return kOk;
}

int ensure_slots(Workspace* ws, int N, int nwaves) {
  size_t need = batched_scratch_doubles(N, nwaves);
  if (need <= ws->slots_count) return kOk;
  if (ws->d_slots) ws->retired.push_back(ws->d_slots);
  ws->d_slots = nullptr;
  ws->slots_count = 0;
  EDB_TRY(dmalloc(&ws->d_slots, need));
  ws->slots_count = need;
  return kOk;
}

void ws_free(Workspace* w) {
  dfree(&w->d_slots);
  dfree(&w->d_dup);
  dfree(&w->d_part);
  for (void* p : w->retired) (void)hipFree(p);
  w->retired.clear();
  w->slots_count = 0;
  w->part_slabs = 0;
}

int ensure_moment_table(edb_qcalc* q) {
  if (q->d_mom) return kOk;
  std::vector<double> tab;
  int E = 0;
  EDB_TRY(build_moment_table(q->N, q->L, &tab, &E));
  double* d = nullptr;
  EDB_TRY(dmalloc(&d, tab.size()));
  if (hipMemcpy(d, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("moment table: copy to the device failed");
    dfree(&d);
    return kHipError;
  }
  q->d_mom = d;  // published only once filled
  q->mom_E = E;
  return kOk;
}

int batched_count(const edb_qcalc* q) {
  if (q->schedule == 1)
    return q->nblocks_quad > 0 ? q->nblocks_quad : cu_count() * 16 / quad_waves_per_block(q->N);
  return q->nwaves > 0 ? q->nwaves : cu_count() * 4 * batched_waves_per_simd(q->N);
}

int batched_shape(edb_qcalc* q, Workspace* ws, int groups, int hermitian, PrefixEntry** pe, int* count) {
  EDB_TRY(get_prefix(q, groups, hermitian ? 1 : 0, pe));
  int c = batched_count(q);
  if (c > (*pe)->total) c = (int)(*pe)->total;
  EDB_TRY(ensure_slots(ws, q->N, q->schedule == 1 ? c * quad_waves_per_block(q->N) : c));
  *count = c;
  return kOk;
}

// Batched Q on device arrays already in the kernel layout.
int batched_device(edb_qcalc* q, Workspace* ws, int batch, int ld, int hermitian, double* d_q,
                   const float* d_fin, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  EDB_REQUIRE(q->d_pair, "calQ_batched: genB_batched has not run");
  EDB_REQUIRE(batch >= 1 && ld % 64 == 0 && batch <= ld, "calQ_batched: bad batch / ld");
  const int groups = (batch + 63) / 64;
  PrefixEntry* pe = nullptr;
  int count = 0;
  EDB_TRY(batched_shape(q, ws, groups, hermitian, &pe, &count));
  if (q->schedule == 1)
    return launch_calq_quad_impl(q->N, q->d_pair, q->nblocks, q->d_sched, q->d_rowinfo, q->d_quadinfo, q->d_rowidx,
                                 q->d_ublk, pe->d_prefix, pe->nitems, pe->total, d_fin, ld, groups,
                                 hermitian ? 1 : 0, d_q, ws->d_slots, count, s, ev0, ev1);
  return launch_calq_batched_impl(q->N, q->d_pair, q->nblocks, q->d_sched, q->d_rowinfo, pe->d_prefix, pe->nitems,
                                  pe->total, d_fin, ld, groups, hermitian ? 1 : 0, d_q, ws->d_slots,
                                  count, s, ev0, ev1);
}

// l-slabs of the on-the-fly plain kernel, each at least 1024 l.  k3-block
// layout: about eight waves per SIMD.  Orbit layout: up to 32, so that many
// waves work on each orbit's rows of F at once (the waves of one XCD then span
// a few orbits), at the cost of 16 bytes per mode and slab of partial sums.
int otf_slabs(const edb_qcalc* q) {
  long ns;
  if (otf_orbit_layout(q->N)) {
    ns = 32;
  } else {
    const long nkb = (long)q->g.M * q->g.M * ((q->g.M + kOtfK - 1) / kOtfK);
    ns = (cu_count() * 32L + nkb - 1) / nkb;
  }
  ns = std::min(ns, std::max(1L, q->g.NQ / 1024));
  return (int)std::max(1L, std::min(ns, 4096L));
}

// l-slabs of the plain kernel that takes beta from `src`.
int plain_slabs(const edb_qcalc* q, int src) {
  if (src == kPlainCompressed) return otf_slabs(q);
  if (src == kPlainFP32 && q->nslab32 > 0) return q->nslab32;
  return q->nslab;
}

// The difference table and room for `slabs` slab partial sums.  A partial
// buffer that is too small is retired (kernels in flight may still read it)
// and replaced.
int ensure_plain_scratch(const edb_qcalc* q, Workspace* ws, int slabs) {
  if (!ws->d_dup) {
    long W = 2 * q->g.M - 1;
    EDB_TRY(dmalloc(&ws->d_dup, (size_t)(W * W * W * 2)));
  }
  if (ws->part_slabs < slabs) {
    if (ws->d_part) ws->retired.push_back(ws->d_part);
    ws->d_part = nullptr;
    ws->part_slabs = 0;
    EDB_TRY(dmalloc(&ws->d_part, (size_t)slabs * q->g.NQ * 2));
    ws->part_slabs = slabs;
  }
  return kOk;
}

// Plain FP64 Q of one column: d_fc FP64 centred [M^3] with stride fc_stride
// (complex elements), result centred into d_q with stride q_stride.
int plain_column(edb_qcalc* q, Workspace* ws, const double* d_fc, int fc_stride, double* d_q, int q_stride,
                 hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, int src) {
  if (src == kPlainFP32) EDB_REQUIRE(q->d_plain32, "calQ_plain32: genB_plain32 has not run");
  else if (src == kPlainCompressed) EDB_REQUIRE(q->d_ftab, "calQ_compressed: genB_compressed has not run");
  else EDB_REQUIRE(q->d_plain, "calQ: genB has not run");
  const int ns = plain_slabs(q, src);
  EDB_TRY(ensure_plain_scratch(q, ws, ns));
  EDB_TRY(launch_build_dup(q->N, d_fc, fc_stride, ws->d_dup, s));
  if (ev0) EDB_HIP_TRY(hipEventRecord(ev0, s));
  if (src == kPlainFP32)
    EDB_TRY(launch_calq_plain32(q->N, q->d_plain32, ws->d_dup, d_fc, fc_stride, ws->d_part, ns, q->pkern32, s));
  else if (src == kPlainCompressed)
    EDB_TRY(launch_calq_plain_otf(q->N, q->tables.rval, q->d_ftab, q->d_sb, ws->d_dup, d_fc, fc_stride,
                                  q->d_kgrp, q->n_kgrp, ws->d_part, ns, s));
  else
    EDB_TRY(launch_calq_plain64(q->N, q->d_plain, ws->d_dup, d_fc, fc_stride, ws->d_part, ns, q->pkern, s));
  if (ev1) EDB_HIP_TRY(hipEventRecord(ev1, s));
  EDB_TRY(launch_reduce_plain(q->N, ws->d_part, ns, d_q, q_stride, s));
  return kOk;
}

// Runs one genB build; when any step of it fails, drains the stream and frees
// every buffer that build owns, so a later call starts again from nothing
// instead of finding a half-built (or never-filled) array and reporting success.
template <typename Build, typename Release>
int build_or_release(edb_qcalc* q, Build build, Release release) {
  const int st = build();
  if (st != EDB_OK) {
    (void)hipStreamSynchronize(q->stream);
    (void)hipGetLastError();
    release();
  }
  return st;
}

}  // namespace edb

extern "C" {

const char* edb_last_error(void) { return edb::last_error(); }

int edb_version(int* major, int* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return EDB_OK;
}

int edb_device_count(int* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return EDB_OK;
}

int edb_sync(void) {
  EDB_HIP_TRY(hipDeviceSynchronize());
  return EDB_OK;
}

// ------------------------------------------------------------ QCalculator

int edb_qcalc_create(int N, int alpha, double L, edb_qcalc** out) {
  EDB_REQUIRE(out, "edb_qcalc_create: null output");
  *out = nullptr;
  // batched kernels exist up to N = 33; the plain paths (stored or compressed beta) run up to 64
  EDB_REQUIRE(N >= 3 && N <= kMaxN, "edb_qcalc_create: N must be in [3, 64]");
  EDB_REQUIRE(alpha == 0 || alpha == 1, "edb_qcalc_create: alpha must be 0 or 1");
  EDB_REQUIRE(L > 0. && std::isfinite(L), "edb_qcalc_create: L must be positive");
  edb_qcalc* q = new edb_qcalc();
  q->N = N;
  q->alpha = alpha;
  q->L = L;
  q->g = Geometry(N);
  q->ntile = (int)((q->g.NQ + 63) / 64);
  // as many l-slabs as the default stream kernel keeps resident at once (one
  // round of workgroups, no tail): 32 at N = 24, 0.19 ms per Q-evaluation
  const int occ = plain_blocks_per_cu(N, q->pkern, false);
  q->nslab = occ > 0 ? cu_count() * occ / q->ntile : (cu_count() * 24 + q->ntile - 1) / q->ntile;
  q->nslab = (int)std::max(1L, std::min<long>({(long)q->nslab, 4096L, q->g.NQ}));
  // the FP32-beta stream is half as long and wants ~4x the waves in flight (measured at N = 24),
  // capped at 64 slabs like the FP64 stream so small grids do not drown in partial sums
  q->nslab32 = (int)std::max(1L, std::min<long>({64L, q->g.NQ, (cu_count() * 48L + q->ntile - 1) / q->ntile}));
  // measured default: workgroup quads for the k3-split grids (N >= 27), where
  // they beat stream-K rows by ~10%; rows elsewhere (the autotuner times both)
  q->schedule = batched_ksplit(N) > 1 ? EDB_SCHED_QUADS : EDB_SCHED_ROWS;
  hipError_t e = hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    delete q;
    return EDB_EHIP;
  }
  *out = q;
  return EDB_OK;
}

int edb_qcalc_close_filter(edb_qcalc* q) {
  EDB_REQUIRE(q, "null handle");
  EDB_REQUIRE(!q->d_plain && !q->d_plain32 && !q->d_pair && !q->d_ftab, "closeFilter must precede genB");
  q->filter = 0;
  if (q->tables_ready) {
    beta_tables_free(&q->tables);
    q->tables_ready = false;
  }
  return EDB_OK;
}

int edb_qcalc_genB(edb_qcalc* q) {
  EDB_REQUIRE(q, "null handle");
  std::lock_guard<std::mutex> lk(q->mu);
  if (q->d_plain) return EDB_OK;
  return build_or_release(
      q,
      [&]() -> int {
        EDB_TRY(ensure_tables(q));
        EDB_TRY(dmalloc(&q->d_plain, (size_t)q->ntile * q->g.NQ * 64));
        EDB_TRY(launch_genb_plain64(q->tables, q->d_plain, q->ntile, q->stream));
        EDB_HIP_TRY(hipStreamSynchronize(q->stream));
        return EDB_OK;
      },
      [&] { dfree(&q->d_plain); });
}

int edb_qcalc_genB_plain32(edb_qcalc* q) {
  EDB_REQUIRE(q, "null handle");
  std::lock_guard<std::mutex> lk(q->mu);
  if (q->d_plain32) return EDB_OK;
  return build_or_release(
      q,
      [&]() -> int {
        EDB_TRY(ensure_tables(q));
        EDB_TRY(dmalloc(&q->d_plain32, (size_t)q->ntile * q->g.NQ * 64));
        EDB_TRY(launch_genb_plain32(q->tables, q->d_plain32, q->ntile, q->stream));
        EDB_HIP_TRY(hipStreamSynchronize(q->stream));
        return EDB_OK;
      },
      [&] { dfree(&q->d_plain32); });
}

int edb_qcalc_genB_compressed(edb_qcalc* q) {
  EDB_REQUIRE(q, "null handle");
  std::lock_guard<std::mutex> lk(q->mu);
  if (q->d_ftab) return EDB_OK;  // set only by a build that finished
  return build_or_release(
      q,
      [&]() -> int {
        EDB_TRY(ensure_tables(q));
        long R, Rh;
        compressed_beta_dims(q->N, &R, &Rh);
        std::vector<int> grp;
        const int ngroups = build_orbit_groups(q->N, &grp);
        EDB_REQUIRE(ngroups >= 1, "genB_compressed: no mode groups for this grid");
        EDB_TRY(dmalloc(&q->d_ftab, (size_t)(R * Rh)));
        EDB_TRY(dmalloc(&q->d_sb, (size_t)q->g.NQ * 2));
        EDB_TRY(dmalloc(&q->d_kgrp, grp.size()));
        EDB_HIP_TRY(hipMemcpy(q->d_kgrp, grp.data(), sizeof(int) * grp.size(), hipMemcpyHostToDevice));
        q->n_kgrp = ngroups;
        EDB_TRY(launch_genb_compressed(q->tables, q->d_ftab, q->d_sb, q->stream));
        EDB_HIP_TRY(hipStreamSynchronize(q->stream));
        return EDB_OK;
      },
      [&] {
        dfree(&q->d_ftab);
        dfree(&q->d_sb);
        dfree(&q->d_kgrp);
        q->n_kgrp = 0;
      });
}

int edb_qcalc_compressed_bytes(const edb_qcalc* q, long* bytes) {
  EDB_REQUIRE(q && bytes, "compressed_bytes: null argument");
  long R, Rh;
  compressed_beta_dims(q->N, &R, &Rh);
  *bytes = (R * Rh + 2 * q->g.NQ) * (long)sizeof(double);
  return EDB_OK;
}

int edb_qcalc_genB_batched(edb_qcalc* q) {
  EDB_REQUIRE(q, "null handle");
  std::lock_guard<std::mutex> lk(q->mu);
  if (q->d_pair) return EDB_OK;  // set only by a build that finished
  if (!batched_supported(q->N)) {
    set_error("genB_batched: no compiled batched kernel for N = " + std::to_string(q->N));
    return EDB_EUNSUPPORTED;
  }
  float* pair = nullptr;
  return build_or_release(
      q,
      [&]() -> int {
        EDB_TRY(ensure_tables(q));
        std::vector<uint32_t> sched;
        build_pair_plan(q->N, &q->rowinfo, &sched, &q->nrows_half);
        q->nblocks = (long)sched.size();
        EDB_TRY(dmalloc(&q->d_rowinfo, q->rowinfo.size()));
        EDB_TRY(dmalloc(&q->d_sched, sched.size()));
        EDB_HIP_TRY(hipMemcpy(q->d_rowinfo, q->rowinfo.data(), sizeof(int4) * q->rowinfo.size(),
                              hipMemcpyHostToDevice));
        EDB_HIP_TRY(
            hipMemcpy(q->d_sched, sched.data(), sizeof(uint32_t) * sched.size(), hipMemcpyHostToDevice));
        std::vector<int> rowidx, ublk;
        EDB_TRY(build_quad_plan(q->N, q->rowinfo, sched, q->nrows_half, &q->quadinfo, &rowidx, &ublk,
                                &q->nquads_half));
        EDB_TRY(dmalloc(&q->d_quadinfo, q->quadinfo.size()));
        EDB_TRY(dmalloc(&q->d_rowidx, rowidx.size()));
        EDB_TRY(dmalloc(&q->d_ublk, ublk.size()));
        EDB_HIP_TRY(hipMemcpy(q->d_quadinfo, q->quadinfo.data(), sizeof(int4) * q->quadinfo.size(),
                              hipMemcpyHostToDevice));
        EDB_HIP_TRY(
            hipMemcpy(q->d_rowidx, rowidx.data(), sizeof(int) * rowidx.size(), hipMemcpyHostToDevice));
        EDB_HIP_TRY(hipMemcpy(q->d_ublk, ublk.data(), sizeof(int) * ublk.size(), hipMemcpyHostToDevice));
        // the paired array is published in d_pair only once it is filled
        EDB_TRY(dmalloc(&pair, (size_t)q->nblocks * q->g.M * q->g.M));
        EDB_TRY(launch_genb_pair32(q->tables, (const int*)q->d_rowinfo, q->d_sched,
                                   (int)q->rowinfo.size(), q->nblocks, pair, q->stream));
        EDB_HIP_TRY(hipStreamSynchronize(q->stream));
        q->d_pair = pair;
        return EDB_OK;
      },
      [&] {
        dfree(&pair);
        dfree(&q->d_rowinfo);
        dfree(&q->d_sched);
        dfree(&q->d_quadinfo);
        dfree(&q->d_rowidx);
        dfree(&q->d_ublk);
        q->rowinfo.clear();
        q->quadinfo.clear();
        q->nblocks = 0;
        q->nrows_half = 0;
        q->nquads_half = 0;
      });
}

namespace {
int calq_device_impl(edb_qcalc* q, double* d_qout, const double* d_fin, void* stream, int src) {
  EDB_REQUIRE(q && d_qout && d_fin, "calQ_device: null argument");
  hipStream_t s = stream ? (hipStream_t)stream : q->stream;
  double *fc = nullptr, *qc = nullptr;
  EDB_TRY(dmalloc(&fc, (size_t)q->g.NQ * 2));
  int st = dmalloc(&qc, (size_t)q->g.NQ * 2);
  if (st == kOk) st = launch_pack(q->N, 1, 1, false, fc, d_fin, nullptr, 0., s);
  if (st == kOk) st = plain_column(q, &q->ws, fc, 1, qc, 1, s, nullptr, nullptr, src);
  if (st == kOk && q->conserve) st = launch_moment_correct(q->d_mom, q->mom_E, 1, 1, qc, s);
  if (st == kOk) st = launch_unpack(q->N, 1, 1, d_qout, qc, s);
  if (st == kOk && hipStreamSynchronize(s) != hipSuccess) {
    set_error("calQ_device: stream synchronisation failed");
    st = kHipError;
  }
  dfree(&fc);
  dfree(&qc);
  return st;
}

int calq_host_impl(edb_qcalc* q, double* qout, const double* fin, int src) {
  EDB_REQUIRE(q && qout && fin, "calQ: null argument");
  size_t bytes = sizeof(double) * 2 * q->g.N3;
  double *din = nullptr, *dout = nullptr;
  EDB_TRY(dmalloc(&din, (size_t)q->g.N3 * 2));
  int st = dmalloc(&dout, (size_t)q->g.N3 * 2);
  if (st == kOk && hipMemcpy(din, fin, bytes, hipMemcpyHostToDevice) != hipSuccess) st = kHipError;
  if (st == kOk) st = calq_device_impl(q, dout, din, q->stream, src);
  if (st == kOk && hipMemcpy(qout, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess) st = kHipError;
  if (st == kHipError && g_err.empty()) set_error("calQ: copy failed");
  dfree(&din);
  dfree(&dout);
  return st;
}
}  // namespace

int edb_qcalc_calQ_device(edb_qcalc* q, double* d_qout, const double* d_fin, void* stream) {
  return calq_device_impl(q, d_qout, d_fin, stream, kPlainFP64);
}

int edb_qcalc_calQ(edb_qcalc* q, double* qout, const double* fin) {
  return calq_host_impl(q, qout, fin, kPlainFP64);
}

int edb_qcalc_calQ_plain32(edb_qcalc* q, double* qout, const double* fin) {
  return calq_host_impl(q, qout, fin, kPlainFP32);
}

int edb_qcalc_calQ_compressed(edb_qcalc* q, double* qout, const double* fin) {
  return calq_host_impl(q, qout, fin, kPlainCompressed);
}

int edb_qcalc_calQ_compressed_device(edb_qcalc* q, double* d_qout, const double* d_fin, void* stream) {
  return calq_device_impl(q, d_qout, d_fin, stream, kPlainCompressed);
}

int edb_qcalc_calQ_batched_device(edb_qcalc* q, int batch, int ld, int hermitian, double* d_qout,
                                  const float* d_fin, void* stream) {
  EDB_REQUIRE(q && d_qout && d_fin, "calQ_batched_device: null argument");
  hipStream_t s = stream ? (hipStream_t)stream : q->stream;
  EDB_TRY(batched_device(q, &q->ws, batch, ld, hermitian, d_qout, d_fin, s));
  if (q->conserve) EDB_TRY(launch_moment_correct(q->d_mom, q->mom_E, batch, ld, d_qout, s));
  return EDB_OK;
}

int edb_qcalc_calQ_batched(edb_qcalc* q, int batch, int hermitian, double* qout, const double* fin) {
  EDB_REQUIRE(q && qout && fin && batch >= 1, "calQ_batched: bad argument");
  const int ld = pad64(batch);
  const long N3 = q->g.N3, NQ = q->g.NQ;
  hipStream_t s = q->stream;
  double *dhost = nullptr, *dgrid = nullptr, *dq = nullptr;
  float* din = nullptr;
  int st = dmalloc(&dhost, (size_t)N3 * batch * 2);
  if (st == kOk) st = dmalloc(&dgrid, (size_t)N3 * ld * 2);
  if (st == kOk) st = dmalloc(&dq, (size_t)NQ * ld * 2);
  if (st == kOk) st = dmalloc(&din, (size_t)NQ * ld * 2);
  if (st == kOk &&
      hipMemcpy(dhost, fin, sizeof(double) * 2 * N3 * batch, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("calQ_batched: host-to-device copy failed");
    st = kHipError;
  }
  if (st == kOk) st = launch_gather_cols(N3, batch, ld, dhost, dgrid, s);
  if (st == kOk) st = launch_pack(q->N, batch, ld, true, din, dgrid, nullptr, 0., s);
  if (st == kOk) st = batched_device(q, &q->ws, batch, ld, hermitian, dq, din, s);
  if (st == kOk && q->conserve) st = launch_moment_correct(q->d_mom, q->mom_E, batch, ld, dq, s);
  if (st == kOk) st = launch_unpack(q->N, batch, ld, dgrid, dq, s);
  if (st == kOk) st = launch_scatter_cols(N3, batch, ld, dgrid, dhost, s);
  if (st == kOk && hipStreamSynchronize(s) != hipSuccess) {
    set_error("calQ_batched: stream synchronisation failed");
    st = kHipError;
  }
  if (st == kOk &&
      hipMemcpy(qout, dhost, sizeof(double) * 2 * N3 * batch, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("calQ_batched: device-to-host copy failed");
    st = kHipError;
  }
  dfree(&dhost);
  dfree(&dgrid);
  dfree(&dq);
  dfree(&din);
  return st;
}

int edb_qcalc_copy_B(edb_qcalc* q, double* Bout) {
  EDB_REQUIRE(q && Bout, "copy_B: null argument");
  EDB_REQUIRE(q->g.NQ <= 50000, "copy_B: M^6 too large to copy (M <= 36)");
  EDB_TRY(ensure_tables(q));
  long total = q->g.NQ * q->g.NQ;
  double* d = nullptr;
  EDB_TRY(dmalloc(&d, (size_t)total));
  int st = launch_genb_ref64(q->tables, d, q->stream);
  if (st == kOk && hipMemcpyAsync(Bout, d, sizeof(double) * total, hipMemcpyDeviceToHost, q->stream) !=
                       hipSuccess)
    st = kHipError;
  if (st == kOk && hipStreamSynchronize(q->stream) != hipSuccess) st = kHipError;
  if (st == kHipError) set_error("copy_B: copy failed");
  dfree(&d);
  return st;
}

int edb_qcalc_info(const edb_qcalc* q, int* N, int* M, long* bytes_plain64, long* bytes_pair32) {
  EDB_REQUIRE(q, "null handle");
  if (N) *N = q->N;
  if (M) *M = q->g.M;
  if (bytes_plain64) *bytes_plain64 = (long)q->ntile * q->g.NQ * 64 * 8;
  if (bytes_pair32) {
    if (q->d_pair) {
      *bytes_pair32 = q->nblocks * q->g.M * q->g.M * 4;
    } else {
      std::vector<int4> ri;
      std::vector<uint32_t> sc;
      int nh = 0;
      build_pair_plan(q->N, &ri, &sc, &nh);
      *bytes_pair32 = (long)sc.size() * q->g.M * q->g.M * 4;
    }
  }
  return EDB_OK;
}

// Device buffers for timing the batched kernel on random coefficients.
struct BatchedTiming {
  int ld = 0;
  float* din = nullptr;
  double* dq = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~BatchedTiming() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(&din);
    dfree(&dq);
  }
};

static int batched_timing_init(const edb_qcalc* q, int batch, BatchedTiming* T) {
  T->ld = pad64(batch);
  const size_t n = (size_t)q->g.NQ * T->ld * 2;
  EDB_TRY(dmalloc(&T->din, n));
  EDB_TRY(dmalloc(&T->dq, n));
  std::vector<float> h(n);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  for (auto& v : h) {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    v = (float)((double)(x >> 11) / 9007199254740992.0 - 0.5);
  }
  if (hipMemcpy(T->din, h.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipEventCreate(&T->e0) != hipSuccess || hipEventCreate(&T->e1) != hipSuccess) {
    set_error("time_batched: HIP failure");
    return kHipError;
  }
  return kOk;
}

static int batched_events_ms(edb_qcalc* q, BatchedTiming* T, int batch, int hermitian, int launches, double* ms) {
  int st = kOk;
  if (hipEventRecord(T->e0, q->stream) != hipSuccess) st = kHipError;
  for (int r = 0; st == kOk && r < launches; ++r)
    st = batched_device(q, &q->ws, batch, T->ld, hermitian, T->dq, T->din, q->stream);
  if (st == kOk && hipEventRecord(T->e1, q->stream) != hipSuccess) st = kHipError;
  if (st == kOk && hipEventSynchronize(T->e1) != hipSuccess) st = kHipError;
  float el = 0.f;
  if (st == kOk && hipEventElapsedTime(&el, T->e0, T->e1) != hipSuccess) st = kHipError;
  if (st == kOk) *ms = el / launches;
  if (st == kHipError) set_error("time_batched: HIP failure");
  return st;
}

// Mean time per call over `reps` back-to-back calls, measured after a warm-up
// of about 40 ms of back-to-back calls.  The device's clock drops over the
// first ~10-20 launches that follow an idle period and then settles (at N=24
// with 256 instances a call takes ~1.18 ms first, up to ~1.5 ms a few calls
// later, ~1.13 ms once settled); the warm-up makes the timing that of a
// sustained run of the kernel.
static int batched_steady_ms(edb_qcalc* q, BatchedTiming* T, int batch, int hermitian, int reps, double* ms) {
  double est = 0.;
  EDB_TRY(batched_events_ms(q, T, batch, hermitian, 1, &est));  // first call: plans, prefix tables
  EDB_TRY(batched_events_ms(q, T, batch, hermitian, 1, &est));
  const int warm = (int)std::min(400.0, std::max(2.0, std::ceil(40.0 / std::max(est, 1e-3))));
  for (int r = 0; r < warm; ++r)
    EDB_TRY(batched_device(q, &q->ws, batch, T->ld, hermitian, T->dq, T->din, q->stream));
  return batched_events_ms(q, T, batch, hermitian, reps, ms);
}

int edb_qcalc_time_batched(edb_qcalc* q, int batch, int hermitian, int reps, double* ms) {
  EDB_REQUIRE(q && ms && reps >= 1 && batch >= 1, "time_batched: bad argument");
  BatchedTiming T;
  EDB_TRY(batched_timing_init(q, batch, &T));
  return batched_steady_ms(q, &T, batch, hermitian, reps, ms);
}

static int time_plain_src(edb_qcalc* q, int reps, double* ms, int src) {
  EDB_REQUIRE(q && ms && reps >= 1, "time_plain: bad argument");
  const long NQ = q->g.NQ;
  double *fc = nullptr, *dq = nullptr;
  EDB_TRY(dmalloc(&fc, (size_t)NQ * 2));
  int st = dmalloc(&dq, (size_t)NQ * 2);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == kOk) {
    std::vector<double> h((size_t)NQ * 2);
    for (size_t i = 0; i < h.size(); ++i) h[i] = std::sin(0.37 * (double)i);
    if (hipMemcpy(fc, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
      st = kHipError;
  }
  if (st == kOk && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = kHipError;
  // first call (scratch allocation), then ~20 ms of back-to-back calls so that
  // the timed ones run at the device's sustained clock
  if (st == kOk) st = plain_column(q, &q->ws, fc, 1, dq, 1, q->stream, nullptr, nullptr, src);
  if (st == kOk && hipEventRecord(e0, q->stream) != hipSuccess) st = kHipError;
  if (st == kOk) st = plain_column(q, &q->ws, fc, 1, dq, 1, q->stream, nullptr, nullptr, src);
  if (st == kOk && hipEventRecord(e1, q->stream) != hipSuccess) st = kHipError;
  if (st == kOk && hipEventSynchronize(e1) != hipSuccess) st = kHipError;
  float est = 0.f;
  if (st == kOk && hipEventElapsedTime(&est, e0, e1) != hipSuccess) st = kHipError;
  const int warm = (int)std::min(400.0, std::max(2.0, std::ceil(20.0 / std::max((double)est, 1e-3))));
  for (int r = 0; st == kOk && r < warm; ++r)
    st = plain_column(q, &q->ws, fc, 1, dq, 1, q->stream, nullptr, nullptr, src);
  if (st == kOk && hipEventRecord(e0, q->stream) != hipSuccess) st = kHipError;
  for (int r = 0; st == kOk && r < reps; ++r)
    st = plain_column(q, &q->ws, fc, 1, dq, 1, q->stream, nullptr, nullptr, src);
  if (st == kOk && hipEventRecord(e1, q->stream) != hipSuccess) st = kHipError;
  if (st == kOk && hipEventSynchronize(e1) != hipSuccess) st = kHipError;
  float el = 0.f;
  if (st == kOk && hipEventElapsedTime(&el, e0, e1) != hipSuccess) st = kHipError;
  if (st == kOk) *ms = el / reps;
  if (st == kHipError) set_error("time_plain: HIP failure");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(&fc);
  dfree(&dq);
  return st;
}

int edb_qcalc_time_plain(edb_qcalc* q, int reps, double* ms) { return time_plain_src(q, reps, ms, kPlainFP64); }

int edb_qcalc_time_compressed(edb_qcalc* q, int reps, double* ms) {
  return time_plain_src(q, reps, ms, kPlainCompressed);
}

int edb_qcalc_tune(edb_qcalc* q, int batched_waves, int plain_slabs) {
  EDB_REQUIRE(q, "null handle");
  EDB_REQUIRE(batched_waves >= 0 && batched_waves <= (1 << 20), "tune: batched_waves out of range");
  EDB_REQUIRE(plain_slabs >= 0 && plain_slabs <= 4096, "tune: plain_slabs out of range");
  std::lock_guard<std::mutex> lk(q->mu);
  q->nwaves = batched_waves;
  if (plain_slabs > 0) q->nslab = (int)std::min<long>(plain_slabs, q->g.NQ);  // d_part grows on use
  return EDB_OK;
}

int edb_qcalc_tune_plain(edb_qcalc* q, int fp32, int kernel, int slabs) {
  EDB_REQUIRE(q, "null handle");
  EDB_REQUIRE(kernel >= -1 && kernel < kPlainKernels, "tune_plain: unknown kernel variant");
  EDB_REQUIRE(slabs >= 0 && slabs <= 4096, "tune_plain: slabs out of range");
  std::lock_guard<std::mutex> lk(q->mu);
  int& kern = fp32 ? q->pkern32 : q->pkern;
  int& ns = fp32 ? q->nslab32 : q->nslab;
  if (kernel >= 0) kern = kernel;
  if (slabs > 0) ns = (int)std::min<long>(slabs, q->g.NQ);  // the partial sums grow on use
  return EDB_OK;
}

int edb_qcalc_plain_schedule(const edb_qcalc* q, int fp32, int* kernel, int* slabs) {
  EDB_REQUIRE(q, "null handle");
  if (kernel) *kernel = fp32 ? q->pkern32 : q->pkern;
  if (slabs) *slabs = fp32 ? (q->nslab32 > 0 ? q->nslab32 : q->nslab) : q->nslab;
  return EDB_OK;
}

int edb_qcalc_tune_batched(edb_qcalc* q, int schedule, int count) {
  EDB_REQUIRE(q, "null handle");
  EDB_REQUIRE(schedule == EDB_SCHED_ROWS || schedule == EDB_SCHED_QUADS, "tune_batched: unknown schedule");
  EDB_REQUIRE(count >= 0 && count <= (1 << 20), "tune_batched: count out of range");
  std::lock_guard<std::mutex> lk(q->mu);
  q->schedule = schedule;
  if (schedule == EDB_SCHED_QUADS) q->nblocks_quad = count;
  else q->nwaves = count;
  return EDB_OK;
}

int edb_qcalc_batched_schedule(const edb_qcalc* q, int* schedule, int* count) {
  EDB_REQUIRE(q, "null handle");
  if (schedule) *schedule = q->schedule;
  if (count) *count = q->schedule == EDB_SCHED_QUADS ? q->nblocks_quad : q->nwaves;
  return EDB_OK;
}

int edb_qcalc_autotune(edb_qcalc* q, int batch, int* batched_waves, int* plain_slabs) {
  EDB_REQUIRE(q, "null handle");
  const int cu = cu_count();
  if (q->d_pair && batch >= 1) {
    BatchedTiming T;
    EDB_TRY(batched_timing_init(q, batch, &T));
    double best = 1e30;
    int bsched = 0, bcount = 0;
    for (int per_cu : {4, 6, 8, 12, 16}) {  // stream-K rows: persistent waves per CU
      q->schedule = EDB_SCHED_ROWS;
      q->nwaves = cu * per_cu;
      double ms = 0.;
      EDB_TRY(batched_steady_ms(q, &T, batch, 1, 5, &ms));
      if (ms < best) {
        best = ms;
        bsched = EDB_SCHED_ROWS;
        bcount = q->nwaves;
      }
    }
    const int wpb = quad_waves_per_block(q->N);
    for (int per_cu : {8, 16, 24, 32}) {  // quads: persistent workgroups, as waves per CU
      q->schedule = EDB_SCHED_QUADS;
      q->nblocks_quad = std::max(1, cu * per_cu / wpb);
      double ms = 0.;
      EDB_TRY(batched_steady_ms(q, &T, batch, 1, 5, &ms));
      if (ms < best) {
        best = ms;
        bsched = EDB_SCHED_QUADS;
        bcount = q->nblocks_quad;
      }
    }
    q->schedule = bsched;
    if (bsched == EDB_SCHED_QUADS) q->nblocks_quad = bcount;
    else q->nwaves = bcount;
  }
  // l-slab candidates of a stream variant: waves per CU from 8 to 64, and the
  // counts whose workgroups all fit on the device at once (one round, no
  // tail of late workgroups) at full, 3/4 and 1/2 of its occupancy
  auto slab_candidates = [&](int kern, bool fp32, std::vector<int>* out) {
    out->clear();
    for (int waves_per_cu : {8, 12, 16, 20, 24, 28, 32, 40, 48, 64})
      out->push_back((cu * waves_per_cu + q->ntile - 1) / q->ntile);
    const int occ = plain_blocks_per_cu(q->N, kern, fp32);
    for (int f4 : {4, 3, 2})
      if (occ > 0) out->push_back(cu * occ * f4 / 4 / q->ntile);
    for (int& v : *out) v = (int)std::min<long>(q->g.NQ, std::max(1, std::min(4096, v)));
    std::sort(out->begin(), out->end());
    out->erase(std::unique(out->begin(), out->end()), out->end());
  };
  if (q->d_plain) {  // stream variant x l-slabs (waves in flight); the best 3 timed again
    std::vector<std::pair<double, std::pair<int, int>>> cand;
    std::vector<int> slabs;
    for (int kern = 0; kern < kPlainKernels; ++kern) {
      q->pkern = kern;
      slab_candidates(kern, false, &slabs);
      for (int ns : slabs) {
        q->nslab = ns;
        double ms = 0.;
        EDB_TRY(edb_qcalc_time_plain(q, 5, &ms));
        cand.push_back({ms, {kern, q->nslab}});
      }
    }
    std::sort(cand.begin(), cand.end());
    double best = 1e30;
    int bk = cand[0].second.first, bs = cand[0].second.second;
    for (size_t i = 0; i < cand.size() && i < 3; ++i) {
      q->pkern = cand[i].second.first;
      q->nslab = cand[i].second.second;
      double ms = 0.;
      EDB_TRY(edb_qcalc_time_plain(q, 20, &ms));
      if (ms < best) {
        best = ms;
        bk = q->pkern;
        bs = q->nslab;
      }
    }
    q->pkern = bk;
    q->nslab = bs;
  }
  if (q->d_plain32) {  // the FP32-beta stream is shorter and wants more slabs in flight
    double best = 1e30;
    int bs = 0, bk = q->pkern32;
    for (int kern = 0; kern < kPlainKernels; ++kern) {
      q->pkern32 = kern;
      for (int waves_per_cu : {16, 32, 48, 64, 96}) {
        q->nslab32 = std::max(1, (int)std::min<long>({4096L, q->g.NQ / 16, (cu * waves_per_cu + q->ntile - 1) / q->ntile}));
        double ms = 0.;
        EDB_TRY(time_plain_src(q, 3, &ms, kPlainFP32));
        if (ms < best) {
          best = ms;
          bs = q->nslab32;
          bk = kern;
        }
      }
    }
    q->nslab32 = bs;
    q->pkern32 = bk;
  }
  if (batched_waves)
    *batched_waves = q->schedule == EDB_SCHED_QUADS ? q->nblocks_quad * quad_waves_per_block(q->N) : q->nwaves;
  if (plain_slabs) *plain_slabs = q->nslab;
  return EDB_OK;
}

int edb_qcalc_set_conservative(edb_qcalc* q, int on) {
  EDB_REQUIRE(q, "null handle");
  std::lock_guard<std::mutex> lk(q->mu);
  if (on) EDB_TRY(ensure_moment_table(q));
  q->conserve = on ? 1 : 0;
  return EDB_OK;
}

int edb_qcalc_destroy(edb_qcalc* q) {
  if (!q) return EDB_OK;
  if (q->stream) (void)hipStreamSynchronize(q->stream);
  dfree(&q->d_mom);
  dfree(&q->d_plain32);
  dfree(&q->d_ftab);
  dfree(&q->d_sb);
  dfree(&q->d_kgrp);
  dfree(&q->d_plain);
  ws_free(&q->ws);
  dfree(&q->d_pair);
  dfree(&q->d_rowinfo);
  dfree(&q->d_sched);
  dfree(&q->d_quadinfo);
  dfree(&q->d_rowidx);
  dfree(&q->d_ublk);
  for (auto& kv : q->prefixes) dfree(&kv.second.d_prefix);
  if (q->tables_ready) beta_tables_free(&q->tables);
  if (q->stream) (void)hipStreamDestroy(q->stream);
  delete q;
  return EDB_OK;
}

// ----------------------------------------------------------------- fft3d

static int fft_host(int Nx, int Ny, int Nz, const double* in, double* out, int sign, double scale) {
  EDB_REQUIRE(in && out, "fft3d: null argument");
  EDB_REQUIRE(Nx >= 1 && Ny >= 1 && Nz >= 1 && Nx <= kMaxN && Ny <= kMaxN && Nz <= kMaxN,
              "fft3d: lengths must be in [1, 64]");
  long n = (long)Nx * Ny * Nz;
  const int ld = 64;
  double *dh = nullptr, *dg = nullptr, *dt = nullptr;
  int st = dmalloc(&dh, (size_t)n * 2);
  if (st == kOk) st = dmalloc(&dg, (size_t)n * ld * 2);
  if (st == kOk) st = dmalloc(&dt, (size_t)n * ld * 2);
  hipStream_t s = nullptr;
  if (st == kOk && hipMemcpy(dh, in, sizeof(double) * 2 * n, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("fft3d: copy failed");
    st = kHipError;
  }
  if (st == kOk) st = launch_gather_cols(n, 1, ld, dh, dg, s);
  if (st == kOk) st = launch_fft3d(Nx, Ny, Nz, 1, ld, sign, scale, dg, dg, dt, s);
  if (st == kOk) st = launch_scatter_cols(n, 1, ld, dg, dh, s);
  if (st == kOk && hipMemcpy(out, dh, sizeof(double) * 2 * n, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("fft3d: copy failed");
    st = kHipError;
  }
  dfree(&dh);
  dfree(&dg);
  dfree(&dt);
  return st;
}

int edb_fft3d_fft(int Nx, int Ny, int Nz, const double* re, double* fre) {
  return fft_host(Nx, Ny, Nz, re, fre, -1, 1.0);
}

int edb_fft3d_ifft(int Nx, int Ny, int Nz, double* re, const double* fre) {
  return fft_host(Nx, Ny, Nz, fre, re, +1, 1. / (double)((long)Nx * Ny * Nz));
}

int edb_fft3d_batched_device(int Nx, int Ny, int Nz, int batch, int ld, int inverse, const double* d_in,
                             double* d_out, void* stream) {
  EDB_REQUIRE(d_in && d_out, "fft3d_batched_device: null argument");
  double scale = inverse ? 1. / (double)((long)Nx * Ny * Nz) : 1.0;
  return launch_fft3d(Nx, Ny, Nz, batch, ld, inverse ? +1 : -1, scale, d_in, d_out, nullptr,
                      (hipStream_t)stream);
}

// ------------------------------------------------------------ EntropyFix

int edb_efix_init(edb_efix_state* st, long Nt, double eps) {
  EDB_REQUIRE(st && Nt >= 1, "efix_init: bad argument");
  st->Nt = Nt;
  st->eps = eps;
  st->average = 1.0;  // EntropyFix.h:20
  st->average_flag = 0;
  return EDB_OK;
}

static int efix_run(edb_efix_state* st, int mode, long Nf, double* f_host, bool write_back, double aim,
                    double* result) {
  EDB_REQUIRE(st && f_host, "EntropyFix: null argument");
  EDB_REQUIRE(Nf == st->Nt, "EntropyFix: Nf != Nt (the reference asserts Nf == Nt)");
  double *df = nullptr, *ds = nullptr, *dt = nullptr;
  int s = dmalloc(&df, (size_t)Nf);
  if (s == kOk) s = dmalloc(&ds, 3);
  if (s == kOk && mode != 0) s = dmalloc(&dt, (size_t)Nf);
  double hst[3] = {st->average, (double)st->average_flag, 0.};
  if (s == kOk && (hipMemcpy(df, f_host, sizeof(double) * Nf, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(ds, hst, sizeof(hst), hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("EntropyFix: copy failed");
    s = kHipError;
  }
  if (s == kOk) s = launch_efix(mode, Nf, df, st->eps, aim, ds, mode != 0 ? dt : df, nullptr);
  if (s == kOk && hipMemcpy(hst, ds, sizeof(hst), hipMemcpyDeviceToHost) != hipSuccess) s = kHipError;
  if (s == kOk && write_back &&
      hipMemcpy(f_host, df, sizeof(double) * Nf, hipMemcpyDeviceToHost) != hipSuccess)
    s = kHipError;
  if (s == kOk) {
    st->average = hst[0];
    st->average_flag = hst[1] != 0.;
    if (result) *result = hst[2];
  } else if (s == kHipError) {
    set_error("EntropyFix: HIP failure");
  }
  dfree(&df);
  dfree(&ds);
  dfree(&dt);
  return s;
}

int edb_efix_set_average(edb_efix_state* st, long Nf, const double* f) {
  return efix_run(st, 0, Nf, const_cast<double*>(f), false, 0., nullptr);
}

int edb_efix_entropy_cal(edb_efix_state* st, long Nf, const double* f, double* H) {
  EDB_REQUIRE(H, "EntropyCal: null output");
  return efix_run(st, 1, Nf, const_cast<double*>(f), false, 0., H);
}

int edb_efix_fixsol(edb_efix_state* st, long Nf, double* f, double aim) {
  return efix_run(st, 2, Nf, f, true, aim, nullptr);
}

}  // extern "C"

// ----------------------------------------------------------------- solver

struct edb_solver {
  edb_solver_params p{};
  edb_qcalc* qc = nullptr;
  bool own_qc = false;
  // Q scratch of this solver's stream (not the qcalc's, which other solvers
  // sharing the collision kernel may be using at the same time)
  edb::Workspace ws;
  Geometry g;
  int ld = 0;
  hipStream_t stream = nullptr;
  double *Fh = nullptr, *Fr = nullptr, *Ft = nullptr, *S1 = nullptr, *S2 = nullptr, *Fold = nullptr;
  double* Q[4] = {nullptr, nullptr, nullptr, nullptr};
  void* in = nullptr;
  double *H0 = nullptr, *avg = nullptr;
  int *avg_flag = nullptr, *fixed = nullptr;
  double *hist_H = nullptr, *hist_Hfix = nullptr, *hist_mean = nullptr, *hist_mean_fix = nullptr;
  int* hist_fixed = nullptr;
  int* d_rec = nullptr;
  double *d_out = nullptr;
  double t = 0.;
  int ref_step = 0;     // the driver's own step counter
  int taken = 0;        // steps taken in total
  int dufort_phase = 0; // 0 start-up RK2 steps, 1 DuFort-Frankel
  bool ready = false;
  // graphs keyed by (phase, eligible)
  bool use_graph = false;
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  // Q-kernel timing
  bool qtiming = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t ev_used = 0;
};

namespace {

bool batched(const edb_solver* s) { return s->p.precision == EDB_PREC_FP32_BATCHED; }
int plain_source(const edb_solver* s) {
  return s->p.precision == EDB_PREC_FP32_PLAIN
             ? kPlainFP32
             : (s->p.precision == EDB_PREC_FP64_COMPRESSED ? kPlainCompressed : kPlainFP64);
}

int next_events(edb_solver* s, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (!s->qtiming || s->use_graph) return kOk;
  if (s->ev_used == s->ev.size()) {
    hipEvent_t a, b;
    EDB_HIP_TRY(hipEventCreate(&a));
    EDB_HIP_TRY(hipEventCreate(&b));
    s->ev.push_back({a, b});
  }
  *e0 = s->ev[s->ev_used].first;
  *e1 = s->ev[s->ev_used].second;
  s->ev_used++;
  return kOk;
}

// Q of the stage state X (+ c0*Qprev) into Qout (centred, [M^3][ld]).
int qeval(edb_solver* s, const double* X, const double* Qprev, double c0, double* Qout) {
  const int N = s->p.N, batch = s->p.batch, ld = s->ld;
  EDB_TRY(launch_pack(N, batch, ld, batched(s), s->in, X, Qprev, c0, s->stream));
  if (batched(s)) {
    hipEvent_t e0, e1;
    EDB_TRY(next_events(s, &e0, &e1));
    EDB_TRY(batched_device(s->qc, &s->ws, batch, ld, s->p.hermitian, Qout, (const float*)s->in, s->stream, e0,
                           e1));
  } else {
    for (int i = 0; i < batch; ++i) {
      hipEvent_t e0, e1;
      EDB_TRY(next_events(s, &e0, &e1));
      EDB_TRY(plain_column(s->qc, &s->ws, (const double*)s->in + 2 * i, ld, Qout + 2 * i, ld, s->stream, e0, e1,
                           plain_source(s)));
    }
  }
  if (s->p.conserve) EDB_TRY(launch_moment_correct(s->qc->d_mom, s->qc->mom_E, batch, ld, Qout, s->stream));
  return kOk;
}

int combine(edb_solver* s, int op, double* out, const double* X, const double* Y, const double* Q0,
            const double* Q1, const double* Q2, const double* Q3, double c0, double c1 = 0.,
            double c2 = 0.) {
  return launch_combine(op, s->p.N, s->p.batch, s->ld, out, X, Y, Q0, Q1, Q2, Q3, c0, c1, c2,
                        s->stream);
}

// ifft, calEntropy, entropy fix (and the forward transform of fixed
// instances in applied mode): the tail of every driver's loop body.
int post_step(edb_solver* s, int eligible) {
  const int N = s->p.N, batch = s->p.batch, ld = s->ld;
  const long N3 = s->g.N3;
  EDB_TRY(launch_fft3d(N, N, N, batch, ld, +1, 1. / (double)N3, s->Fh, s->Fr, s->Ft, s->stream));
  EntropyFixArgs a;
  a.N3 = N3;
  a.ld = ld;
  a.F = (double2*)s->Fr;
  a.W = s->S1;
  a.part = s->S1 + N3 * ld;  // second half of S1 (W uses the first)
  a.H0 = s->H0;
  a.avg = s->avg;
  a.avg_flag = s->avg_flag;
  a.fixed = s->fixed;
  a.hist_H = s->hist_H;
  a.hist_Hfix = s->hist_Hfix;
  a.hist_fixed = s->hist_fixed;
  a.Fh = (const double2*)s->Fh;
  a.hist_mean = s->hist_mean;
  a.hist_mean_fix = s->hist_mean_fix;
  a.rec_ptr = s->d_rec;
  a.hist_cap = s->p.history;
  a.eligible = eligible;
  a.fix_mode = s->p.fix_mode;
  a.eps = 1.0e-10;
  double R = s->p.L;
  a.vol = std::pow(2 * R, 3.);
  a.vol_over_N3 = std::pow(2 * R, 3.) / (double)N3;
  EDB_TRY(launch_entropy_fix(a, batch, s->stream));
  if (s->p.fix_mode == EDB_FIX_APPLIED && eligible) {
    // only the fixed instances are transformed back (the others keep Fh)
    EDB_TRY(launch_fft3d(N, N, N, batch, ld, -1, 1.0, s->Fr, s->Ft, s->S2, s->stream, s->fixed));
    EDB_TRY(launch_select_cols(N3, batch, ld, s->fixed, s->Ft, s->Fh, s->stream));
  }
  EDB_TRY(launch_bump(s->d_rec, s->stream));
  return kOk;
}

// One pass of the driver's loop body (without the host-side t update).
int step_body(edb_solver* s, int phase, int eligible) {
  const double dt = s->p.dt;
  double** Q = s->Q;
  switch (s->p.scheme) {
    case EDB_SCHEME_FEULER:  // BKW.h:220-222
      EDB_TRY(qeval(s, s->Fh, nullptr, 0., Q[0]));
      EDB_TRY(combine(s, kOpAxpy, s->Fh, s->Fh, nullptr, Q[0], nullptr, nullptr, nullptr, dt));
      break;
    case EDB_SCHEME_CFE: {  // BKW.h:281-283
      double c = (1. - std::exp(-s->p.beta * dt)) / s->p.beta;
      EDB_TRY(qeval(s, s->Fh, nullptr, 0., Q[0]));
      EDB_TRY(combine(s, kOpAxpy, s->Fh, s->Fh, nullptr, Q[0], nullptr, nullptr, nullptr, c));
      break;
    }
    case EDB_SCHEME_RK2:  // BKW.h:341-346
      EDB_TRY(qeval(s, s->Fh, nullptr, 0., Q[0]));
      EDB_TRY(qeval(s, s->Fh, Q[0], 0.5 * dt, Q[1]));
      EDB_TRY(combine(s, kOpAxpy, s->Fh, s->Fh, nullptr, Q[1], nullptr, nullptr, nullptr, dt));
      break;
    case EDB_SCHEME_SSP3:  // BKW.h:416-424
      EDB_TRY(qeval(s, s->Fh, nullptr, 0., Q[0]));
      EDB_TRY(combine(s, kOpAxpy, s->S1, s->Fh, nullptr, Q[0], nullptr, nullptr, nullptr, dt));
      EDB_TRY(qeval(s, s->S1, nullptr, 0., Q[1]));
      EDB_TRY(combine(s, kOpSSP3b, s->S2, s->Fh, s->S1, Q[1], nullptr, nullptr, nullptr, 0.25 * dt));
      EDB_TRY(qeval(s, s->S2, nullptr, 0., Q[2]));
      EDB_TRY(combine(s, kOpSSP3c, s->Fh, s->Fh, s->S2, Q[2], nullptr, nullptr, nullptr, 2. / 3. * dt));
      break;
    case EDB_SCHEME_RK4:  // BKW.h:482-494
      EDB_TRY(qeval(s, s->Fh, nullptr, 0., Q[0]));
      EDB_TRY(qeval(s, s->Fh, Q[0], 0.5 * dt, Q[1]));
      EDB_TRY(qeval(s, s->Fh, Q[1], 0.5 * dt, Q[2]));
      EDB_TRY(qeval(s, s->Fh, Q[2], dt, Q[3]));
      EDB_TRY(combine(s, kOpRK4, s->Fh, s->Fh, nullptr, Q[0], Q[1], Q[2], Q[3], dt / 6.));
      break;
    case EDB_SCHEME_DUFORT:
      if (phase == 0) {  // start-up RK2 with step dt, BKW.h:551-556
        EDB_TRY(qeval(s, s->Fh, nullptr, 0., Q[0]));
        EDB_TRY(qeval(s, s->Fh, Q[0], 0.5 * dt, Q[1]));
        EDB_TRY(combine(s, kOpAxpy, s->Fh, s->Fh, nullptr, Q[1], nullptr, nullptr, nullptr, dt));
      } else {  // BKW.h:595-598, 630
        const double b = s->p.beta;
        EDB_TRY(qeval(s, s->Fh, nullptr, 0., Q[0]));
        EDB_TRY(combine(s, kOpCopy, s->S1, s->Fh, nullptr, nullptr, nullptr, nullptr, nullptr, 0.));
        EDB_TRY(combine(s, kOpDufort, s->Fh, s->S1, s->Fold, Q[0], nullptr, nullptr, nullptr,
                        1. / (1. + b * dt), dt, b * dt));
        EDB_TRY(combine(s, kOpCopy, s->Fold, s->S1, nullptr, nullptr, nullptr, nullptr, nullptr, 0.));
      }
      break;
    default:
      set_error("unknown scheme");
      return kInvalid;
  }
  return post_step(s, eligible);
}

int run_one(edb_solver* s) {
  const int scheme = s->p.scheme;
  // fix eligibility of this step, per driver
  bool fix_on = s->p.fix_mode != EDB_FIX_NONE;
  if (s->p.fix_mode == EDB_FIX_DIAGNOSTIC && (scheme == EDB_SCHEME_RK2 || scheme == EDB_SCHEME_DUFORT))
    fix_on = false;  // `if(0)` at BKW.h:358, 568, 610
  int eligible = fix_on && (scheme == EDB_SCHEME_FEULER ? s->ref_step > 1 : s->ref_step > 0);
  int phase = (scheme == EDB_SCHEME_DUFORT) ? s->dufort_phase : 0;
  if (scheme == EDB_SCHEME_DUFORT && phase == 0) s->t += s->p.dt / 8;
  else s->t += s->p.dt;

  auto key = std::make_pair(phase, eligible);
  if (s->use_graph) {
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      EDB_HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      int st = step_body(s, phase, eligible);
      hipError_t e = hipStreamEndCapture(s->stream, &graph);
      if (st != kOk) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      EDB_HIP_TRY(e);
      EDB_HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
      it = s->graphs.emplace(key, exec).first;
    }
    EDB_HIP_TRY(hipGraphLaunch(it->second, s->stream));
  } else {
    EDB_TRY(step_body(s, phase, eligible));
  }
  s->ref_step++;
  s->taken++;
  if (scheme == EDB_SCHEME_DUFORT && phase == 0 && s->ref_step >= 8) {  // while(step < 8), BKW.h:587
    s->dufort_phase = 1;
    s->ref_step = 0;
  }
  return kOk;
}

// Everything a step allocates lazily, allocated now, so that a step can be
// captured into a HIP graph (no allocation during capture).
int prepare_buffers(edb_solver* s) {
  EDB_TRY(fft_prepare(s->p.N));
  if (s->p.conserve) EDB_TRY(ensure_moment_table(s->qc));
  if (batched(s)) {
    const int groups = (s->p.batch + 63) / 64;
    PrefixEntry* pe = nullptr;
    int count = 0;
    EDB_TRY(batched_shape(s->qc, &s->ws, groups, s->p.hermitian ? 1 : 0, &pe, &count));
  } else {
    EDB_TRY(ensure_plain_scratch(s->qc, &s->ws, plain_slabs(s->qc, plain_source(s))));
  }
  return kOk;
}

int init_common(edb_solver* s, double T0) {
  const int batch = s->p.batch, ld = s->ld;
  EDB_TRY(prepare_buffers(s));
  s->t = T0;
  s->taken = 0;
  s->dufort_phase = 0;
  s->ref_step = (s->p.scheme == EDB_SCHEME_FEULER) ? 1 : 0;
  EDB_TRY(launch_fill(s->avg, ld, 1.0, s->stream));  // Averagef = 1.0, EntropyFix.h:20
  EDB_TRY(launch_fill((double*)s->avg_flag, (ld + 1) / 2, 0., s->stream));
  EDB_TRY(launch_fill((double*)s->d_rec, 1, 0., s->stream));
  if (s->p.scheme == EDB_SCHEME_FEULER) {
    // entropy0 = calEntropy() of the initial F, BKW.h:214; no fix, no record
    EDB_TRY(launch_fft3d(s->p.N, s->p.N, s->p.N, batch, ld, +1, 1. / (double)s->g.N3, s->Fh, s->Fr,
                         s->Ft, s->stream));
    EntropyFixArgs a;
    a.N3 = s->g.N3;
    a.ld = ld;
    a.F = (double2*)s->Fr;
    a.W = s->S1;
    a.part = s->S1 + s->g.N3 * ld;
    a.H0 = s->H0;
    a.avg = s->avg;
    a.avg_flag = s->avg_flag;
    a.eligible = 0;
    a.fix_mode = 0;
    a.vol = std::pow(2 * s->p.L, 3.);
    a.vol_over_N3 = a.vol / (double)s->g.N3;
    EDB_TRY(launch_entropy_fix(a, batch, s->stream));
  } else {
    EDB_TRY(launch_fill(s->H0, ld, 1., s->stream));  // entropy0 = 1., BKW.h:478
  }
  if (s->p.scheme == EDB_SCHEME_DUFORT)  // F_old = Fhat, BKW.h:539
    EDB_TRY(combine(s, kOpCopy, s->Fold, s->Fh, nullptr, nullptr, nullptr, nullptr, nullptr, 0.));
  EDB_HIP_TRY(hipStreamSynchronize(s->stream));
  s->ready = true;
  return kOk;
}

}  // namespace

extern "C" {

int edb_solver_create(const edb_solver_params* p, edb_qcalc* shared, edb_solver** out) {
  EDB_REQUIRE(p && out, "solver_create: null argument");
  *out = nullptr;
  EDB_REQUIRE(p->batch >= 1 && p->batch <= 65536, "solver_create: batch must be in [1, 65536]");
  EDB_REQUIRE(p->precision == EDB_PREC_FP64_PLAIN || p->precision == EDB_PREC_FP32_BATCHED ||
                  p->precision == EDB_PREC_FP32_PLAIN || p->precision == EDB_PREC_FP64_COMPRESSED,
              "solver_create: unknown precision");
  EDB_REQUIRE(p->scheme >= 0 && p->scheme <= 5, "solver_create: unknown scheme");
  EDB_REQUIRE(p->fix_mode >= 0 && p->fix_mode <= 2, "solver_create: unknown fix mode");
  EDB_REQUIRE(p->dt > 0. && std::isfinite(p->dt), "solver_create: dt must be positive");
  EDB_REQUIRE((p->scheme != EDB_SCHEME_CFE && p->scheme != EDB_SCHEME_DUFORT) || p->beta > 0.,
              "solver_create: CFE/DUFORT need beta > 0");
  EDB_REQUIRE(p->history >= 0, "solver_create: history must be >= 0");
  // the batched kernels exist up to N = 33; the plain precisions (stored or compressed beta, generic
  // DFT passes past 33) up to 64
  if (p->precision == EDB_PREC_FP32_BATCHED)
    EDB_REQUIRE(p->N >= 3 && p->N <= kMaxNCalq, "solver_create: batched precision needs N in [3, 33]");
  else
    EDB_REQUIRE(p->N >= 3 && p->N <= kMaxN, "solver_create: N must be in [3, 64]");
  if (shared)
    EDB_REQUIRE(shared->N == p->N && shared->alpha == p->alpha && shared->L == p->L &&
                    shared->filter == (p->filter ? 1 : 0),
                "solver_create: shared qcalc has different (N, alpha, L, filter)");
  edb_solver* s = new edb_solver();
  s->p = *p;
  s->g = Geometry(p->N);
  s->ld = pad64(p->batch);
  int st = kOk;
  if (shared) {
    s->qc = shared;
  } else {
    st = edb_qcalc_create(p->N, p->alpha, p->L, &s->qc);
    s->own_qc = st == kOk;
    if (st == kOk && !p->filter) st = edb_qcalc_close_filter(s->qc);
  }
  if (st == kOk)
    st = batched(s) ? edb_qcalc_genB_batched(s->qc)
                    : (plain_source(s) == kPlainFP32
                           ? edb_qcalc_genB_plain32(s->qc)
                           : (plain_source(s) == kPlainCompressed ? edb_qcalc_genB_compressed(s->qc)
                                                                  : edb_qcalc_genB(s->qc)));
  if (st == kOk && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("solver_create: stream creation failed");
    st = kHipError;
  }
  const size_t grid = (size_t)s->g.N3 * s->ld * 2, modes = (size_t)s->g.NQ * s->ld * 2;
  if (st == kOk) st = dmalloc(&s->Fh, grid);
  if (st == kOk) st = dmalloc(&s->Fr, grid);
  if (st == kOk) st = dmalloc(&s->Ft, grid);
  if (st == kOk) st = dmalloc(&s->S1, grid);
  if (st == kOk) st = dmalloc(&s->S2, grid);
  if (st == kOk && p->scheme == EDB_SCHEME_DUFORT) st = dmalloc(&s->Fold, grid);
  for (int i = 0; i < 4 && st == kOk; ++i) st = dmalloc(&s->Q[i], modes);
  if (st == kOk) st = dmalloc((double**)&s->in, modes);
  if (st == kOk) st = dmalloc(&s->H0, (size_t)s->ld);
  if (st == kOk) st = dmalloc(&s->avg, (size_t)s->ld);
  if (st == kOk) st = dmalloc(&s->avg_flag, (size_t)s->ld + 2);
  if (st == kOk) st = dmalloc(&s->fixed, (size_t)s->ld);
  if (st == kOk) st = dmalloc((double**)&s->d_rec, 1);
  if (st == kOk) st = dmalloc(&s->d_out, (size_t)s->ld * 5);
  if (st == kOk && p->history > 0) {
    st = dmalloc(&s->hist_H, (size_t)p->history * s->ld);
    if (st == kOk) st = dmalloc(&s->hist_Hfix, (size_t)p->history * s->ld);
    if (st == kOk) st = dmalloc(&s->hist_fixed, (size_t)p->history * s->ld);
    if (st == kOk) st = dmalloc(&s->hist_mean, (size_t)p->history * s->ld);
    if (st == kOk) st = dmalloc(&s->hist_mean_fix, (size_t)p->history * s->ld);
  }
  if (st == kOk) st = launch_fill(s->S2, grid, 0., s->stream);
  if (st == kOk) st = launch_fill((double*)s->fixed, (s->ld + 1) / 2, 0., s->stream);
  if (st != kOk) {
    std::string msg = g_err;
    edb_solver_destroy(s);
    set_error(msg);
    return st;
  }
  *out = s;
  return EDB_OK;
}

int edb_solver_set_initial_real(edb_solver* s, double T0, const double* F) {
  EDB_REQUIRE(s && F, "set_initial_real: null argument");
  const long N3 = s->g.N3;
  double* dh = nullptr;
  EDB_TRY(dmalloc(&dh, (size_t)N3 * s->p.batch));
  int st = kOk;
  if (hipMemcpy(dh, F, sizeof(double) * N3 * s->p.batch, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("set_initial_real: copy failed");
    st = kHipError;
  }
  if (st == kOk) st = launch_real_to_complex_cols(N3, s->p.batch, s->ld, dh, s->Fr, s->stream);
  if (st == kOk)
    st = launch_fft3d(s->p.N, s->p.N, s->p.N, s->p.batch, s->ld, -1, 1.0, s->Fr, s->Fh, s->Ft, s->stream);
  if (st == kOk && hipStreamSynchronize(s->stream) != hipSuccess) st = kHipError;
  dfree(&dh);
  if (st == kOk) st = init_common(s, T0);
  return st;
}

int edb_solver_set_initial_spectral(edb_solver* s, double T0, const double* Fhat) {
  EDB_REQUIRE(s && Fhat, "set_initial_spectral: null argument");
  const long N3 = s->g.N3;
  double* dh = nullptr;
  EDB_TRY(dmalloc(&dh, (size_t)N3 * s->p.batch * 2));
  int st = kOk;
  if (hipMemcpy(dh, Fhat, sizeof(double) * 2 * N3 * s->p.batch, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("set_initial_spectral: copy failed");
    st = kHipError;
  }
  if (st == kOk) st = launch_gather_cols(N3, s->p.batch, s->ld, dh, s->Fh, s->stream);
  if (st == kOk && hipStreamSynchronize(s->stream) != hipSuccess) st = kHipError;
  dfree(&dh);
  if (st == kOk) st = init_common(s, T0);
  return st;
}

int edb_solver_step(edb_solver* s, int nsteps) {
  EDB_REQUIRE(s && nsteps >= 0, "solver_step: bad argument");
  if (!s->ready) {
    set_error("solver_step: initial data not set");
    return EDB_ENOTREADY;
  }
  for (int i = 0; i < nsteps; ++i) EDB_TRY(run_one(s));
  return EDB_OK;
}

int edb_solver_run(edb_solver* s, double TEnd, int* nsteps) {
  EDB_REQUIRE(s, "solver_run: null handle");
  if (!s->ready) {
    set_error("solver_run: initial data not set");
    return EDB_ENOTREADY;
  }
  int n = 0;
  while (s->p.scheme == EDB_SCHEME_DUFORT && s->dufort_phase == 0) {  // while(step < 8), BKW.h:587
    EDB_TRY(run_one(s));
    ++n;
  }
  do {
    EDB_TRY(run_one(s));
    ++n;
  } while (s->t < TEnd);
  if (nsteps) *nsteps = n;
  return EDB_OK;
}

int edb_solver_use_graph(edb_solver* s, int enable) {
  EDB_REQUIRE(s, "null handle");
  s->use_graph = enable != 0;
  return EDB_OK;
}

int edb_solver_enable_qtiming(edb_solver* s, int enable) {
  EDB_REQUIRE(s, "null handle");
  s->qtiming = enable != 0;
  s->ev_used = 0;
  return EDB_OK;
}

int edb_solver_qkernel_time(edb_solver* s, double* ms_mean, long* launches) {
  EDB_REQUIRE(s, "null handle");
  EDB_HIP_TRY(hipStreamSynchronize(s->stream));
  double tot = 0.;
  for (size_t i = 0; i < s->ev_used; ++i) {
    float el = 0.f;
    EDB_HIP_TRY(hipEventElapsedTime(&el, s->ev[i].first, s->ev[i].second));
    tot += el;
  }
  if (ms_mean) *ms_mean = s->ev_used ? tot / (double)s->ev_used : 0.;
  if (launches) *launches = (long)s->ev_used;
  s->ev_used = 0;
  return EDB_OK;
}

int edb_solver_time(const edb_solver* s, double* t, int* steps) {
  EDB_REQUIRE(s, "null handle");
  if (t) *t = s->t;
  if (steps) *steps = s->taken;
  return EDB_OK;
}

int edb_solver_get_spectral(edb_solver* s, double* Fhat) {
  EDB_REQUIRE(s && Fhat, "get_spectral: null argument");
  const long N3 = s->g.N3;
  double* dh = nullptr;
  EDB_TRY(dmalloc(&dh, (size_t)N3 * s->p.batch * 2));
  int st = launch_scatter_cols(N3, s->p.batch, s->ld, s->Fh, dh, s->stream);
  if (st == kOk && hipStreamSynchronize(s->stream) != hipSuccess) st = kHipError;
  if (st == kOk && hipMemcpy(Fhat, dh, sizeof(double) * 2 * N3 * s->p.batch, hipMemcpyDeviceToHost) !=
                       hipSuccess)
    st = kHipError;
  if (st == kHipError) set_error("get_spectral: HIP failure");
  dfree(&dh);
  return st;
}

static int refresh_real(edb_solver* s) {
  return launch_fft3d(s->p.N, s->p.N, s->p.N, s->p.batch, s->ld, +1, 1. / (double)s->g.N3, s->Fh,
                      s->Ft, s->S2, s->stream);
}

int edb_solver_get_real(edb_solver* s, double* F) {
  EDB_REQUIRE(s && F, "get_real: null argument");
  const long N3 = s->g.N3;
  double* dh = nullptr;
  EDB_TRY(dmalloc(&dh, (size_t)N3 * s->p.batch));
  int st = refresh_real(s);
  if (st == kOk) st = launch_complex_to_real_cols(N3, s->p.batch, s->ld, s->Ft, dh, s->stream);
  if (st == kOk && hipStreamSynchronize(s->stream) != hipSuccess) st = kHipError;
  if (st == kOk &&
      hipMemcpy(F, dh, sizeof(double) * N3 * s->p.batch, hipMemcpyDeviceToHost) != hipSuccess)
    st = kHipError;
  if (st == kHipError) set_error("get_real: HIP failure");
  dfree(&dh);
  return st;
}

int edb_solver_entropy_history(edb_solver* s, int steps, double* H, double* Hfix, int* fixed) {
  EDB_REQUIRE(s, "null handle");
  EDB_REQUIRE(steps >= 0 && steps <= s->p.history, "entropy_history: steps exceeds the history kept");
  EDB_HIP_TRY(hipStreamSynchronize(s->stream));
  const int b = s->p.batch, ld = s->ld;
  std::vector<double> h((size_t)steps * ld), hf((size_t)steps * ld);
  std::vector<int> fx((size_t)steps * ld);
  if (steps > 0) {
    EDB_HIP_TRY(hipMemcpy(h.data(), s->hist_H, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    EDB_HIP_TRY(hipMemcpy(hf.data(), s->hist_Hfix, sizeof(double) * hf.size(), hipMemcpyDeviceToHost));
    EDB_HIP_TRY(hipMemcpy(fx.data(), s->hist_fixed, sizeof(int) * fx.size(), hipMemcpyDeviceToHost));
  }
  for (int r = 0; r < steps; ++r)
    for (int i = 0; i < b; ++i) {
      if (H) H[(size_t)r * b + i] = h[(size_t)r * ld + i];
      if (Hfix) Hfix[(size_t)r * b + i] = hf[(size_t)r * ld + i];
      if (fixed) fixed[(size_t)r * b + i] = fx[(size_t)r * ld + i];
    }
  return EDB_OK;
}

int edb_solver_mean_history(edb_solver* s, int steps, double* mean, double* mean_fix) {
  EDB_REQUIRE(s, "null handle");
  EDB_REQUIRE(steps >= 0 && steps <= s->p.history, "mean_history: steps exceeds the history kept");
  EDB_HIP_TRY(hipStreamSynchronize(s->stream));
  const int b = s->p.batch, ld = s->ld;
  std::vector<double> m((size_t)steps * ld), mf((size_t)steps * ld);
  if (steps > 0) {
    EDB_HIP_TRY(hipMemcpy(m.data(), s->hist_mean, sizeof(double) * m.size(), hipMemcpyDeviceToHost));
    EDB_HIP_TRY(hipMemcpy(mf.data(), s->hist_mean_fix, sizeof(double) * mf.size(), hipMemcpyDeviceToHost));
  }
  for (int r = 0; r < steps; ++r)
    for (int i = 0; i < b; ++i) {
      if (mean) mean[(size_t)r * b + i] = m[(size_t)r * ld + i];
      if (mean_fix) mean_fix[(size_t)r * b + i] = mf[(size_t)r * ld + i];
    }
  return EDB_OK;
}

int edb_solver_get_driver_real(edb_solver* s, double* F) {
  EDB_REQUIRE(s && F, "get_driver_real: null argument");
  const long N3 = s->g.N3;
  double* dh = nullptr;
  EDB_TRY(dmalloc(&dh, (size_t)N3 * s->p.batch));
  int st = kOk;
  // before the first step the driver's F is the inverse transform of the state
  if (s->taken == 0)
    st = launch_fft3d(s->p.N, s->p.N, s->p.N, s->p.batch, s->ld, +1, 1. / (double)N3, s->Fh, s->Fr,
                      s->S2, s->stream);
  if (st == kOk) st = launch_complex_to_real_cols(N3, s->p.batch, s->ld, s->Fr, dh, s->stream);
  if (st == kOk && hipStreamSynchronize(s->stream) != hipSuccess) st = kHipError;
  if (st == kOk &&
      hipMemcpy(F, dh, sizeof(double) * N3 * s->p.batch, hipMemcpyDeviceToHost) != hipSuccess)
    st = kHipError;
  if (st == kHipError) set_error("get_driver_real: HIP failure");
  dfree(&dh);
  return st;
}

int edb_solver_moments(edb_solver* s, double* out) {
  EDB_REQUIRE(s && out, "moments: null argument");
  EDB_TRY(refresh_real(s));
  EDB_TRY(launch_moments(s->p.N, s->p.batch, s->ld, s->p.L, s->Ft, s->d_out, s->stream));
  EDB_HIP_TRY(hipStreamSynchronize(s->stream));
  EDB_HIP_TRY(hipMemcpy(out, s->d_out, sizeof(double) * 5 * s->p.batch, hipMemcpyDeviceToHost));
  return EDB_OK;
}

int edb_solver_bkw_error(edb_solver* s, double K, double* out) {
  EDB_REQUIRE(s && out, "bkw_error: null argument");
  EDB_REQUIRE(K > 0., "bkw_error: K must be positive");
  EDB_TRY(refresh_real(s));
  EDB_TRY(launch_bkw_error(s->p.N, s->p.batch, s->ld, s->p.L, K, s->Ft, s->d_out, s->stream));
  EDB_HIP_TRY(hipStreamSynchronize(s->stream));
  EDB_HIP_TRY(hipMemcpy(out, s->d_out, sizeof(double) * 2 * s->p.batch, hipMemcpyDeviceToHost));
  return EDB_OK;
}

int edb_solver_destroy(edb_solver* s) {
  if (!s) return EDB_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& e : s->ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  dfree(&s->Fh);
  dfree(&s->Fr);
  dfree(&s->Ft);
  dfree(&s->S1);
  dfree(&s->S2);
  dfree(&s->Fold);
  for (int i = 0; i < 4; ++i) dfree(&s->Q[i]);
  dfree((double**)&s->in);
  dfree(&s->H0);
  dfree(&s->avg);
  dfree(&s->avg_flag);
  dfree(&s->fixed);
  dfree(&s->hist_H);
  dfree(&s->hist_Hfix);
  dfree(&s->hist_fixed);
  dfree(&s->hist_mean);
  dfree(&s->hist_mean_fix);
  dfree((double**)&s->d_rec);
  dfree(&s->d_out);
  ws_free(&s->ws);
  if (s->own_qc) edb_qcalc_destroy(s->qc);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return EDB_OK;
}

}  // extern "C"
