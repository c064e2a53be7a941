// noalloc.hip -- edb_qcalc_calQ_batched_device_noalloc: the batched Q of
// edb_qcalc_calQ_batched_device with every scratch byte in a file-scope
// device array instead of library-allocated memory, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops).
//
// The array holds the work schedule's prefix table (computed on the device in
// closed form) followed by the stream-K partial slots of the batched kernel.
// Each GPU holds its own copy until the process ends; all calls on one GPU
// share it, so calls that overlap on one GPU (on different streams, or while
// an earlier call is still running on another stream) race.  Calls in order
// on one stream are safe: every byte a call reads is written earlier in the
// same call.
#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/edb.h"
#include "capi_internal.hpp"
#include "edb_internal.hpp"
#include "ops.hpp"

using namespace edb;

namespace {

// The schedules' prefix tables in closed form (the same numbers batched_prefix
// and quad_prefix build on the host).  Rows: row r carries w = rowinfo[r].w
// steps per item and rowinfo[r].z = the steps of all rows before it; items
// run (group g, row r, k3 part h), it = g*nunits*ks + r*ks + h, so item it
// starts at g*T + z*ks + h*w with T = (z + w of the last row) * ks.  Quads:
// uoff = quadinfo[qd].w union steps before quad qd, items (group g, quad qd),
// item g*nunits + qd starts at g*U + uoff with U = the union steps of all
// nunits quads.  One thread per entry.
__global__ void prefix_fill_kernel(const int4* __restrict__ info, int quads, int nunits, int per_unit, int ks,
                                   long* __restrict__ prefix) {
  const long nitems = (long)nunits * per_unit;
  const long T = quads ? 0 : ((long)info[nunits - 1].z + info[nunits - 1].w) * ks;
  for (long it = blockIdx.x * (long)blockDim.x + threadIdx.x; it <= nitems; it += (long)gridDim.x * blockDim.x) {
    long v;
    if (quads) {
      const long g = it / nunits;
      const int qd = (int)(it - g * nunits);  // the end entry: g = groups, qd = 0
      v = g * (long)info[nunits].w + info[qd].w;
    } else {
      const long per_g = (long)nunits * ks;  // items of one instance group
      const long g = it / per_g, rem = it - g * per_g;  // the end entry: g = groups, rem = 0
      const int r = (int)(rem / ks), h = (int)(rem - (long)r * ks);
      v = g * T + (long)info[r < nunits ? r : nunits - 1].z * ks + h * info[r < nunits ? r : nunits - 1].w;
    }
    prefix[it] = v;
  }
}

int launch_prefix_fill(const int4* d_info, int quads, int nunits, int per_unit, int ks, long* d_prefix,
                       hipStream_t s) {
  EDB_REQUIRE(nunits >= 1 && per_unit >= 1 && ks >= 1, "prefix_fill: empty schedule");
  const long n = (long)nunits * per_unit + 1;
  const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(prefix_fill_kernel, dim3(blocks), dim3(256), 0, s, d_info, quads, nunits, per_unit, ks,
                     d_prefix);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace

// Scratch of edb_qcalc_calQ_batched_device_noalloc: the schedule's prefix
// table followed by the stream-K partial slots, in file-scope device memory
// (one copy per GPU, held until the process ends).  In host code `scratch`
// is only a handle for hipGetSymbolAddress; kernels get the per-device
// address it returns.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device (with
// that device current) and kept.
int noalloc_scratch(void** out) {
  int dev = 0;
  EDB_HIP_TRY(hipGetDevice(&dev));
  EDB_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "calQ_batched_device_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    EDB_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EDB_REQUIRE(p != nullptr, "calQ_batched_device_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace

extern "C" {

int edb_qcalc_calQ_batched_device_noalloc(edb_qcalc* q, int batch, int ld, int hermitian, double* d_qout,
                                          const float* d_fin, void* stream) {
  EDB_REQUIRE(q && d_qout && d_fin, "calQ_batched_device_noalloc: null argument");
  EDB_REQUIRE(batch >= 1 && ld % 64 == 0 && batch <= ld, "calQ_batched_device_noalloc: bad batch / ld");
  if (!q->d_pair) {
    set_error("calQ_batched_device_noalloc: genB_batched has not run");
    return EDB_ENOTREADY;
  }
  EDB_REQUIRE(!q->conserve || q->d_mom, "calQ_batched_device_noalloc: moment table missing");
  hipStream_t s = stream ? (hipStream_t)stream : q->stream;
  // the launch shape of batched_device, with the prefix table computed on the
  // device in closed form instead of copied from a host table
  const int N = q->N, groups = (batch + 63) / 64;
  const int quads = q->schedule == EDB_SCHED_QUADS ? 1 : 0;
  const int wpb = quads ? quad_waves_per_block(N) : 1;
  const int nunits = quads ? (hermitian ? q->nquads_half : (int)q->quadinfo.size() - 1)
                           : (hermitian ? q->nrows_half : (int)q->rowinfo.size());
  int per_unit = 0;
  long total = 0;
  EDB_REQUIRE(nunits >= 1, "calQ_batched_device_noalloc: empty schedule");
  if (quads) {
    per_unit = groups;
    total = (long)q->quadinfo[nunits].w * groups;
  } else {
    per_unit = batched_ksplit(N) * groups;
    const int4 last = q->rowinfo[nunits - 1];
    total = ((long)last.z + last.w) * per_unit;
  }
  // the closed form needs each unit's start to be the sum of the units before it
  for (int u = 0; u < nunits; ++u) {
    const bool ok = quads ? (u == 0 ? q->quadinfo[0].w == 0 : q->quadinfo[u].w >= q->quadinfo[u - 1].w)
                          : (q->rowinfo[u].z == (u == 0 ? 0 : q->rowinfo[u - 1].z + q->rowinfo[u - 1].w));
    EDB_REQUIRE(ok, "calQ_batched_device_noalloc: schedule offsets are not cumulative");
  }
  const long nitems = (long)nunits * per_unit;
  const size_t prefix_bytes = ((size_t)(nitems + 1) * sizeof(long) + 255) / 256 * 256;
  auto slot_bytes = [&](long count) {
    const size_t w = (size_t)count * wpb;
    return batched_scratch_doubles(N, (int)w) * sizeof(double);
  };
  if (nitems >= (1L << 30) || prefix_bytes + slot_bytes(1) > kNoAllocScratchBytes) {
    set_error("calQ_batched_device_noalloc: batch " + std::to_string(batch) + " at N = " + std::to_string(N) +
              " needs more than the " + std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch");
    return EDB_EINVAL;
  }
  // persistent count as batched_device, capped to what the scratch holds
  long count = std::min<long>(batched_count(q), total);
  const long fit = (long)((kNoAllocScratchBytes - prefix_bytes) / slot_bytes(1));
  count = std::max(1L, std::min(count, fit + 1));
  while (count > 1 && prefix_bytes + slot_bytes(count) > kNoAllocScratchBytes) --count;
  void* base = nullptr;
  EDB_TRY(noalloc_scratch(&base));
  long* d_prefix = (long*)base;
  double* d_slots = (double*)((char*)base + prefix_bytes);
  EDB_TRY(launch_prefix_fill(quads ? q->d_quadinfo : q->d_rowinfo, quads, nunits, per_unit, batched_ksplit(N),
                             d_prefix, s));
  if (quads)
    EDB_TRY(launch_calq_quad_impl(N, q->d_pair, q->nblocks, q->d_sched, q->d_rowinfo, q->d_quadinfo, q->d_rowidx,
                                  q->d_ublk, d_prefix, (int)nitems, total, d_fin, ld, groups, hermitian ? 1 : 0,
                                  d_qout, d_slots, (int)count, s));
  else
    EDB_TRY(launch_calq_batched_impl(N, q->d_pair, q->nblocks, q->d_sched, q->d_rowinfo, d_prefix, (int)nitems,
                                     total, d_fin, ld, groups, hermitian ? 1 : 0, d_qout, d_slots, (int)count, s));
  if (q->conserve) EDB_TRY(launch_moment_correct(q->d_mom, q->mom_E, batch, ld, d_qout, s));
  return EDB_OK;
}


}  // extern "C"
