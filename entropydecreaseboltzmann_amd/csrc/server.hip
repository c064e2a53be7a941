// server.hip -- continuous batching of Boltzmann solves on one GPU.
//
// A request is one instance: (N, alpha, t_end, initial f on the velocity
// grid).  Requests of the same (N, alpha) share a group: one resident
// collision kernel (the FP32 paired layout, built once and cached), a
// device-resident block of `capacity` instance slots (spectral state and the
// per-instance EntropyFix state) and its own HIP stream, so groups of
// different sizes run concurrently on the device.  A group advances all its
// active slots together, RK4 with dt and the applied entropy fix, in chunks
// of steps that end when the next slot finishes; at a chunk boundary the
// finished slots are read out and freed (the slot array is kept compact so the
// batched kernel only ever runs over active instances) and waiting requests
// are admitted into free slots.  A group is created only if its collision
// kernel and slot state fit the HBM budget that remains; a request that could
// never fit is refused at submission.
//
// Each slot's trajectory is exactly the one the same request would follow
// alone (instances never interact; only the FP64 merge order of the batched
// kernel's partial sums depends on the batch), which the tests check.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "../../include/edb.h"
#include "capi_internal.hpp"

using namespace edb;

namespace {

// Longest solve a request may ask for, in steps.
constexpr long kMaxServerSteps = 100000000;

struct Request {
  long id = 0;
  int N = 0, alpha = 0;
  double t_end = 0.;
  double t_final = 0.;  // t after the drivers' last step (accumulated, may exceed t_end)
  int steps = 0;
  std::vector<double> F0;
};

struct Result {
  long id = 0;
  int N = 0, steps = 0;
  double t = 0.;
  std::vector<double> F;
};

struct Slot {
  long id = 0;
  int target = 0, done = 0;
  double t_final = 0.;
};

struct Group {
  int N = 0, alpha = 0, cap = 0, active = 0;
  Geometry g;
  edb_qcalc* qc = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  bool busy = false;
  int chunk = 0;
  double *Fh = nullptr, *Fr = nullptr, *Ft = nullptr, *W = nullptr, *part = nullptr, *Q[4] = {};
  float* in = nullptr;
  double *H0 = nullptr, *avg = nullptr, *out = nullptr, *adm = nullptr;
  int *avg_flag = nullptr, *fixed = nullptr, *steps_done = nullptr, *newflag = nullptr;
  std::vector<Slot> slots;
  std::deque<Request> pending;
  std::vector<double> h_adm;  // host staging of admitted data (kept alive across the async copy)
  std::vector<int> h_flag;
  long bytes = 0;
};

long group_bytes(int N, int cap) {
  Geometry g(N);
  std::vector<int4> ri;
  std::vector<uint32_t> sc;
  int nh = 0;
  build_pair_plan(N, &ri, &sc, &nh);
  long beta = (long)sc.size() * g.M * g.M * 4;
  long per_slot = g.N3 * 16 * 3 + g.N3 * 8 + g.NQ * 16 * 4 + g.NQ * 8 + g.N3 * 8 * 2 + 64;
  return beta + per_slot * cap + (long)entropy_chunks(g.N3, cap) * cap * 8;
}

}  // namespace

struct edb_server {
  edb_server_params p{};
  long used = 0;
  std::map<std::pair<int, int>, std::unique_ptr<Group>> groups;
  std::deque<Result> results;
  long submitted = 0, finished = 0;
};

namespace {

void free_group(Group* g) {
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  dfree(&g->Fh);
  dfree(&g->Fr);
  dfree(&g->Ft);
  dfree(&g->W);
  dfree(&g->part);
  for (auto& q : g->Q) dfree(&q);
  dfree(&g->in);
  dfree(&g->H0);
  dfree(&g->avg);
  dfree(&g->out);
  dfree(&g->adm);
  dfree(&g->avg_flag);
  dfree(&g->fixed);
  dfree(&g->steps_done);
  dfree(&g->newflag);
  if (g->ev) (void)hipEventDestroy(g->ev);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  if (g->qc) edb_qcalc_destroy(g->qc);
}

int make_group(edb_server* s, int N, int alpha, Group** out) {
  auto g = std::make_unique<Group>();
  g->N = N;
  g->alpha = alpha;
  g->g = Geometry(N);
  g->cap = pad64(s->p.max_slots > 0 ? s->p.max_slots : (N >= 27 ? 64 : 256));
  g->bytes = group_bytes(N, g->cap);
  const size_t grid = (size_t)g->g.N3 * g->cap * 2, modes = (size_t)g->g.NQ * g->cap * 2;
  int st = edb_qcalc_create(N, alpha, s->p.L, &g->qc);
  if (st == kOk && !s->p.filter) st = edb_qcalc_close_filter(g->qc);
  if (st == kOk) st = edb_qcalc_genB_batched(g->qc);
  if (st == kOk && hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) st = kHipError;
  if (st == kOk && hipEventCreateWithFlags(&g->ev, hipEventDisableTiming) != hipSuccess) st = kHipError;
  if (st == kOk) st = dmalloc(&g->Fh, grid);
  if (st == kOk) st = dmalloc(&g->Fr, grid);
  if (st == kOk) st = dmalloc(&g->Ft, grid);
  if (st == kOk) st = dmalloc(&g->W, (size_t)g->g.N3 * g->cap);
  if (st == kOk) st = dmalloc(&g->part, (size_t)entropy_chunks(g->g.N3, g->cap) * g->cap);
  for (int i = 0; i < 4 && st == kOk; ++i) st = dmalloc(&g->Q[i], modes);
  if (st == kOk) st = dmalloc(&g->in, modes);
  if (st == kOk) st = dmalloc(&g->H0, (size_t)g->cap);
  if (st == kOk) st = dmalloc(&g->avg, (size_t)g->cap);
  if (st == kOk) st = dmalloc(&g->out, (size_t)g->g.N3 * g->cap);
  if (st == kOk) st = dmalloc(&g->adm, (size_t)g->g.N3 * g->cap);
  if (st == kOk) st = dmalloc(&g->avg_flag, (size_t)g->cap);
  if (st == kOk) st = dmalloc(&g->fixed, (size_t)g->cap);
  if (st == kOk) st = dmalloc(&g->steps_done, (size_t)g->cap);
  if (st == kOk) st = dmalloc(&g->newflag, (size_t)g->cap);
  if (st == kOk) st = fft_prepare(N);
  if (st == kOk && s->p.conserve) st = ensure_moment_table(g->qc);
  if (st == kOk) {
    PrefixEntry* pe = nullptr;
    int count = 0;
    for (int grp = 1; grp <= g->cap / 64 && st == kOk; ++grp)
      st = batched_shape(g->qc, &g->qc->ws, grp, 1, &pe, &count);
  }
  if (st == kOk) st = launch_fill((double*)g->fixed, g->cap / 2, 0., g->stream);
  if (st == kOk && hipStreamSynchronize(g->stream) != hipSuccess) st = kHipError;
  if (st != kOk) {
    std::string msg = edb::last_error();
    free_group(g.get());
    set_error("server: cannot create group: " + msg);
    return st;
  }
  *out = g.get();
  s->used += g->bytes;
  s->groups[{N, alpha}] = std::move(g);
  return kOk;
}

// Admit waiting requests into free slots: stage their real data in `adm`,
// place each in its slot column of Fr, forward-transform and copy only the new
// columns into Fh; reset the new slots' EntropyFix state.
int admit(Group* g) {
  if (g->pending.empty() || g->active >= g->cap) return kOk;
  const long N3 = g->g.N3;
  g->h_adm.resize((size_t)N3 * g->cap);
  g->h_flag.assign(g->cap, 0);
  const int first = g->active;
  while (!g->pending.empty() && g->active < g->cap) {
    Request& r = g->pending.front();
    std::copy(r.F0.begin(), r.F0.end(), g->h_adm.begin() + (long)g->active * N3);
    Slot sl;
    sl.id = r.id;
    sl.target = r.steps;
    sl.t_final = r.t_final;
    g->slots.push_back(sl);
    g->h_flag[g->active] = 1;
    g->active++;
    g->pending.pop_front();
  }
  // The group is idle here (admission happens between chunks), and the next
  // admission waits for the chunk that follows this one, so the host staging
  // buffers stay untouched until these copies have completed.
  EDB_HIP_TRY(hipMemcpyAsync(g->adm + (long)first * N3, g->h_adm.data() + (long)first * N3,
                             sizeof(double) * N3 * (g->active - first), hipMemcpyHostToDevice, g->stream));
  EDB_HIP_TRY(hipMemcpyAsync(g->newflag, g->h_flag.data(), sizeof(int) * g->cap, hipMemcpyHostToDevice,
                             g->stream));
  // new columns of Fr <- staged real data (imaginary part 0), then transform
  EDB_TRY(launch_stage_real(N3, g->active, g->cap, g->newflag, g->adm, g->Fr, g->stream));
  EDB_TRY(launch_fft3d(g->N, g->N, g->N, g->active, g->cap, -1, 1.0, g->Fr, g->Ft, nullptr, g->stream));
  EDB_TRY(launch_select_cols(N3, g->active, g->cap, g->newflag, g->Ft, g->Fh, g->stream));
  EDB_TRY(launch_reset_slots(g->cap, g->newflag, g->H0, g->avg, g->avg_flag, g->steps_done, g->stream));
  return kOk;
}

// One RK4 step (BKW.h:482-494) of the active slots with the applied fix.
int rk4_step(Group* g, const edb_server_params& p) {
  const int N = g->N, b = g->active, ld = g->cap;
  const double dt = p.dt;
  auto qe = [&](const double* X, const double* Qp, double c0, double* Qo) -> int {
    EDB_TRY(launch_pack(N, b, ld, true, g->in, X, Qp, c0, g->stream));
    EDB_TRY(batched_device(g->qc, &g->qc->ws, b, ld, 1, Qo, g->in, g->stream));
    if (p.conserve) EDB_TRY(launch_moment_correct(g->qc->d_mom, g->qc->mom_E, b, ld, Qo, g->stream));
    return kOk;
  };
  EDB_TRY(qe(g->Fh, nullptr, 0., g->Q[0]));
  EDB_TRY(qe(g->Fh, g->Q[0], 0.5 * dt, g->Q[1]));
  EDB_TRY(qe(g->Fh, g->Q[1], 0.5 * dt, g->Q[2]));
  EDB_TRY(qe(g->Fh, g->Q[2], dt, g->Q[3]));
  EDB_TRY(launch_combine(kOpRK4, N, b, ld, g->Fh, g->Fh, nullptr, g->Q[0], g->Q[1], g->Q[2], g->Q[3],
                         dt / 6., 0., 0., g->stream));
  const long N3 = g->g.N3;
  EDB_TRY(launch_fft3d(N, N, N, b, ld, +1, 1. / (double)N3, g->Fh, g->Fr, nullptr, g->stream));
  EntropyFixArgs a;
  a.N3 = N3;
  a.ld = ld;
  a.F = (double2*)g->Fr;
  a.W = g->W;
  a.part = g->part;
  a.H0 = g->H0;
  a.avg = g->avg;
  a.avg_flag = g->avg_flag;
  a.fixed = g->fixed;
  a.eligible = 1;
  a.steps_done = g->steps_done;
  a.fix_mode = p.fix_mode;
  a.vol = std::pow(2 * p.L, 3.);
  a.vol_over_N3 = a.vol / (double)N3;
  EDB_TRY(launch_entropy_fix(a, b, g->stream));
  if (p.fix_mode == EDB_FIX_APPLIED) {
    EDB_TRY(launch_fft3d(N, N, N, b, ld, -1, 1.0, g->Fr, g->Ft, nullptr, g->stream, g->fixed));
    EDB_TRY(launch_select_cols(N3, b, ld, g->fixed, g->Ft, g->Fh, g->stream));
  }
  EDB_TRY(launch_incr(g->steps_done, b, g->stream));
  return kOk;
}

// Read out finished slots and compact the slot array.
int retire(edb_server* s, Group* g) {
  const long N3 = g->g.N3;
  std::vector<int> fin;
  for (int i = 0; i < g->active; ++i)
    if (g->slots[i].done >= g->slots[i].target) fin.push_back(i);
  if (fin.empty()) return kOk;
  for (size_t k = 0; k < fin.size(); ++k)
    EDB_TRY(launch_extract_col(N3, g->cap, fin[k], g->Fr, g->out + (long)k * N3, g->stream));
  std::vector<double> host((size_t)N3 * fin.size());
  EDB_HIP_TRY(hipMemcpyAsync(host.data(), g->out, sizeof(double) * host.size(), hipMemcpyDeviceToHost,
                             g->stream));
  EDB_HIP_TRY(hipStreamSynchronize(g->stream));
  for (size_t k = 0; k < fin.size(); ++k) {
    Result r;
    r.id = g->slots[fin[k]].id;
    r.N = g->N;
    r.steps = g->slots[fin[k]].done;
    r.t = g->slots[fin[k]].t_final;
    r.F.assign(host.begin() + (long)k * N3, host.begin() + (long)(k + 1) * N3);
    s->results.push_back(std::move(r));
    s->finished++;
  }
  // compaction, highest index first
  for (auto it = fin.rbegin(); it != fin.rend(); ++it) {
    int i = *it, last = g->active - 1;
    if (i != last) {
      EDB_TRY(launch_move_col(N3, g->cap, last, i, g->Fh, g->stream));
      EDB_TRY(launch_move_slot_scalars(last, i, g->H0, g->avg, g->avg_flag, g->steps_done, g->stream));
      g->slots[i] = g->slots[last];
    }
    g->slots.pop_back();
    g->active--;
  }
  return kOk;
}

}  // namespace

extern "C" {

int edb_server_create(const edb_server_params* p, edb_server** out) {
  EDB_REQUIRE(p && out, "server_create: null argument");
  *out = nullptr;
  EDB_REQUIRE(p->dt > 0. && p->L > 0., "server_create: dt and L must be positive");
  EDB_REQUIRE(p->fix_mode == EDB_FIX_NONE || p->fix_mode == EDB_FIX_APPLIED,
              "server_create: fix_mode must be NONE or APPLIED");
  EDB_REQUIRE(p->max_slots >= 0 && p->max_slots <= 4096, "server_create: max_slots out of range");
  EDB_REQUIRE(p->max_chunk >= 0, "server_create: max_chunk must be >= 0");
  EDB_REQUIRE(p->max_pending >= 0, "server_create: max_pending must be >= 0");
  edb_server* s = new edb_server();
  s->p = *p;
  if (s->p.max_chunk == 0) s->p.max_chunk = 16;
  *out = s;
  return EDB_OK;
}

int edb_server_submit(edb_server* s, long id, int N, int alpha, double t_end, const double* F0) {
  EDB_REQUIRE(s && F0, "server_submit: null argument");
  EDB_REQUIRE(N >= 3 && N <= kMaxNCalq && batched_supported(N), "server_submit: unsupported N");
  EDB_REQUIRE(alpha == 0 || alpha == 1, "server_submit: alpha must be 0 or 1");
  EDB_REQUIRE(t_end > 0. && std::isfinite(t_end), "server_submit: t_end must be positive");
  // the step count of do { t += dt } while (t < t_end) is found by that loop
  // below; bound it first (the loop would stall once t_end/dt nears 2^53)
  EDB_REQUIRE(t_end / s->p.dt <= (double)kMaxServerSteps,
              "server_submit: t_end/dt exceeds " + std::to_string(kMaxServerSteps) + " steps");
  // a non-finite sample would make the whole solve NaN: refuse it here, before
  // any device work (instances never mix, so it could not spoil the others)
  for (long i = 0, N3 = (long)N * N * N; i < N3; ++i)
    if (!std::isfinite(F0[i])) {
      set_error("server_submit: initial data is not finite at sample " + std::to_string(i));
      return EDB_EINVAL;
    }
  if (s->p.max_pending > 0) {
    long waiting = 0;
    for (auto& kv : s->groups) waiting += (long)kv.second->pending.size();
    if (waiting >= s->p.max_pending) {
      set_error("server_submit: admission limit reached (" + std::to_string(waiting) + " waiting); retry later");
      return EDB_EBUSY;
    }
  }
  auto key = std::make_pair(N, alpha);
  if (!s->groups.count(key)) {
    int cap = pad64(s->p.max_slots > 0 ? s->p.max_slots : (N >= 27 ? 64 : 256));
    long need = group_bytes(N, cap);
    if (s->p.hbm_budget > 0 && need > s->p.hbm_budget) {
      set_error("server_submit: a group for this N needs " + std::to_string(need) +
                " bytes, more than the HBM budget");
      return EDB_ENOMEM;
    }
    if (s->p.hbm_budget > 0 && s->used + need > s->p.hbm_budget) {
      // evict idle groups (their collision kernels) until this one fits
      for (auto it = s->groups.begin(); it != s->groups.end() && s->used + need > s->p.hbm_budget;) {
        Group* g = it->second.get();
        if (g->active == 0 && g->pending.empty() && !g->busy) {
          s->used -= g->bytes;
          free_group(g);
          it = s->groups.erase(it);
        } else {
          ++it;
        }
      }
      if (s->used + need > s->p.hbm_budget) {
        set_error("server_submit: HBM budget exhausted by active groups; retry later");
        return EDB_EBUSY;
      }
    }
    Group* g = nullptr;
    EDB_TRY(make_group(s, N, alpha, &g));
  }
  Group* g = s->groups[key].get();
  Request r;
  r.id = id;
  r.N = N;
  r.alpha = alpha;
  r.t_end = t_end;
  // the drivers' loop: do { t += dt; ... } while (t < TEnd), BKW.h:480-516
  double t = 0.;
  int n = 0;
  do {
    t += s->p.dt;
    ++n;
  } while (t < t_end);
  r.steps = n;
  r.t_final = t;
  r.F0.assign(F0, F0 + g->g.N3);
  g->pending.push_back(std::move(r));
  s->submitted++;
  return EDB_OK;
}

int edb_server_poll(edb_server* s, int* ready) {
  EDB_REQUIRE(s, "server_poll: null handle");
  for (auto& kv : s->groups) {
    Group* g = kv.second.get();
    if (g->busy) {
      hipError_t e = hipEventQuery(g->ev);
      if (e == hipErrorNotReady) continue;
      EDB_HIP_TRY(e);
      g->busy = false;
      for (int i = 0; i < g->active; ++i) g->slots[i].done += g->chunk;
      EDB_TRY(retire(s, g));
    }
    EDB_TRY(admit(g));
    if (g->active == 0) continue;
    int k = s->p.max_chunk;
    for (int i = 0; i < g->active; ++i) k = std::min(k, g->slots[i].target - g->slots[i].done);
    k = std::max(k, 1);
    for (int i = 0; i < k; ++i) EDB_TRY(rk4_step(g, s->p));
    EDB_HIP_TRY(hipEventRecord(g->ev, g->stream));
    g->chunk = k;
    g->busy = true;
  }
  if (ready) *ready = (int)s->results.size();
  return EDB_OK;
}

int edb_server_fetch(edb_server* s, long* id, int* N, int* steps, double* t, double* F, long capacity) {
  EDB_REQUIRE(s && id, "server_fetch: null argument");
  if (s->results.empty()) {
    set_error("server_fetch: no finished result");
    return EDB_ENOTREADY;
  }
  Result& r = s->results.front();
  EDB_REQUIRE(!F || capacity >= (long)r.F.size(), "server_fetch: output buffer too small");
  *id = r.id;
  if (N) *N = r.N;
  if (steps) *steps = r.steps;
  if (t) *t = r.t;
  if (F) std::copy(r.F.begin(), r.F.end(), F);
  s->results.pop_front();
  return EDB_OK;
}

int edb_server_stats(const edb_server* s, long* submitted, long* finished, long* active, long* waiting,
                     long* hbm_used) {
  EDB_REQUIRE(s, "server_stats: null handle");
  long a = 0, w = 0;
  for (auto& kv : s->groups) {
    a += kv.second->active;
    w += (long)kv.second->pending.size();
  }
  if (submitted) *submitted = s->submitted;
  if (finished) *finished = s->finished;
  if (active) *active = a;
  if (waiting) *waiting = w;
  if (hbm_used) *hbm_used = s->used;
  return EDB_OK;
}

int edb_server_destroy(edb_server* s) {
  if (!s) return EDB_OK;
  for (auto& kv : s->groups) free_group(kv.second.get());
  delete s;
  return EDB_OK;
}

}  // extern "C"
