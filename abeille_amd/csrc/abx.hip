// abx.hip - the C ABI (include/abx.h) on gfx950.
//
// Host side of the drop-in boundary: builds and stages the problem tables,
// launches the transport kernels on the problem's own stream, orders the
// fission bank like the reference, and runs the batch / generation drivers
// with banks and tallies resident in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "event.h"
#include "exact_kernels.h"
#include "kernels.h"
#include "tables.h"

namespace {

thread_local std::string g_err;

struct AbxError {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& m) { throw AbxError{code, m}; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(ABX_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

// Grow-only device buffer.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void reserve(size_t b) {
    if (b <= bytes) return;
    if (p) HIPC(hipFree(p));
    p = nullptr;
    HIPC(hipMalloc(&p, b));
    bytes = b;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

int64_t blocks_for(int64_t n, int bs) { return (n + bs - 1) / bs; }

// A kernel is about to read or write n records of `rec` bytes in b: refuse
// to launch it over a buffer that is too small.
void require_records(const DBuf& b, int64_t n, size_t rec, const char* what) {
  if (n > 0 && b.bytes < (size_t)n * rec)
    fail(ABX_E_INVALID, std::string("internal: ") + what + " buffer holds " +
                            std::to_string(b.bytes / rec) + " records, " + std::to_string(n) +
                            " needed");
}

}  // namespace

struct abx_problem {
  HostTables T;
  std::vector<DBuf> arrays;  // uploaded tables
  DevProblem P{};
  DevSources S{};
  hipStream_t stream = nullptr;
  int engine = ABX_ENGINE_HISTORY;
  abx_tuning tuning{256, 0, 1, 0, 0};
  int device = 0;
  // tallies
  DBuf tally_gen, tally_avg, tally_var;
  DBuf tally_more;  // batch engine: generation arrays of concurrently run batches 2, 3, ...
  int tally_g = 0;
  // transport scratch
  DBuf in, sec, stage, stage_parent, stage_daughter, n_stage, daughters, offsets,
      block_sums, block_sums2, ordered, scores, counters, err, err_hist, trace, n_rec, wsum,
      pool_next, blob, gdev, khist_dev;
  EventBuffers ev;
  std::vector<hipEvent_t> gen_events;  // graph runs: 2 per generation
  // cancellation
  bool cancel = false;
  int batch_stage = -1;  // batch kernel: tables staged in LDS pay off (-1 = not yet known)
  DevCancel cm{};
  DBuf ckeys, ckeys2, ctemp, cw;  // sort keys, sort scratch, weights in sorted order
  // exact cancellation: configuration, the parent record of each staged and
  // ordered fission site, the pipeline's buffers, and settings::rng
  int exact = 0;
  DevExact dex{};
  DBuf stage_pinfo, ordered_pinfo, ex_head, ex_offs, ex_bins, ex_smp, ex_fv, ex_fi, ex_cnt,
      ex_state;
  uint64_t grng = 0;
  // driver state
  DBuf bank_a, bank_b;  // abx_particle banks
  size_t bank_n = 0;
  bool run_started = false;
  int rep_chains = 0;       // chains of the replica run that can be continued (0 = none)
  bool use_graphs = true;   // k-eigenvalue runs replay captured generations
  uint64_t run_hist = 0;  // histories_counter
  double run_kcol = 1.0;  // Tallies::kcol() of the previous generation
  // stopping between generations (Simulation::signaled, settings::max_time)
  std::atomic<int> stop_requested{0};
  double max_time = 0.0;  // seconds of wall clock per run, <= 0: none
  std::chrono::steady_clock::time_point run_t0{};
  int run_gens = 0;  // generations completed since the run started
  // per-kernel timing of the transport kernels (HIP events on p->stream)
  bool timing = false;
  double kernel_ms = 0.0;
  int64_t kernel_launches = 0;
  int64_t kernel_items = 0;
  hipEvent_t ke0 = nullptr, ke1 = nullptr;
  // scalar statistics (src/tallies.cpp:268-280)
  int gen = 0;
  // running means and variances of k_col k_abs k_trk k_tot leakage mig
  double s_avg[6] = {0, 0, 0, 0, 0, 0}, s_var[6] = {0, 0, 0, 0, 0, 0};
  // every generation's six values since the run started (the reference's
  // k_col_vec ... mig_vec, inactive generations included)
  std::vector<double> gen_vals[6];
  abx_scores gen_scores{};
  int64_t last_sites = 0;  // ordered fission sites of the last transport (in `ordered`)
  abx_diag run_diag{};     // summed over batch-engine steps since attach
  int64_t run_histories = 0;

  template <class V>
  const typename V::value_type* upload(const V& v) {
    DBuf b;
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(typename V::value_type);
    b.reserve(bytes);
    if (!v.empty()) HIPC(hipMemcpy(b.p, v.data(), v.size() * sizeof(typename V::value_type), hipMemcpyHostToDevice));
    arrays.push_back(b);
    return b.as<typename V::value_type>();
  }

  ~abx_problem() {
    for (auto& b : arrays) b.release();
    DBuf* all[] = {&tally_gen, &tally_avg, &tally_var, &tally_more, &in, &sec, &stage, &stage_parent,
                   &stage_daughter, &n_stage, &daughters, &offsets, &block_sums, &block_sums2,
                   &ordered, &scores, &counters, &err, &err_hist, &trace, &n_rec, &wsum,
                   &bank_a, &bank_b, &ckeys, &ckeys2, &ctemp, &cw, &pool_next, &blob, &gdev,
                   &khist_dev, &stage_pinfo, &ordered_pinfo, &ex_head, &ex_offs, &ex_bins,
                   &ex_smp, &ex_fv, &ex_fi, &ex_cnt, &ex_state};
    for (auto e : gen_events) (void)hipEventDestroy(e);
    for (DBuf* b : all) b->release();
    ev.release();
    if (ke0) (void)hipEventDestroy(ke0);
    if (ke1) (void)hipEventDestroy(ke1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Exclusive scan of n uint32 counts (device) into uint64 offsets; returns total.
uint64_t device_scan(abx_problem* p, const uint32_t* in, uint64_t* out, int64_t n) {
  if (n <= 0) return 0;  // an empty bank: no launch (a grid of 0 blocks is an error)
  const int bs = 1024;
  const int64_t nb = blocks_for(n, bs);
  p->block_sums.reserve((size_t)(nb + 1) * sizeof(uint64_t));
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(bs), 0, p->stream, in, out,
                     p->block_sums.as<uint64_t>(), n);
  HIPC(hipGetLastError());
  // block sums are few (n / 1024): scan them on the host
  std::vector<uint64_t> bsum((size_t)nb);
  HIPC(hipMemcpyAsync(bsum.data(), p->block_sums.p, nb * sizeof(uint64_t), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipStreamSynchronize(p->stream));
  uint64_t acc = 0;
  for (auto& v : bsum) {
    const uint64_t t = v;
    v = acc;
    acc += t;
  }
  p->block_sums2.reserve((size_t)(nb + 1) * sizeof(uint64_t));
  HIPC(hipMemcpyAsync(p->block_sums2.p, bsum.data(), nb * sizeof(uint64_t), hipMemcpyHostToDevice, p->stream));
  hipLaunchKernelGGL(k_add_block_offsets, dim3((unsigned)nb), dim3(bs), 0, p->stream, out,
                     p->block_sums2.as<uint64_t>(), n);
  HIPC(hipGetLastError());
  return acc;
}

// Sort keys are (bin << 32) | bank index, written in index order, with bins
// below nbins and 0xffffffff for no bin.  Sorting on bits [32, end) with
// 2^(end - 32) > nbins orders the bins and puts no-bin last (its low bits
// are all ones, above every bin); the radix sort is stable, so each bin's
// sites stay in bank order without sorting the index bits.
static int bin_key_end_bit(uint64_t nbins) {
  int b = 1;
  while (b < 32 && (1ull << b) <= nbins) b++;
  return 32 + b;
}

// Approximate mesh cancellation of a device bank (in place).
void cancel_device(abx_problem* p, abx_banked* d_bank, int64_t n) {
  if (n <= 1) return;
  p->ckeys.reserve((size_t)n * sizeof(uint64_t));
  p->ckeys2.reserve((size_t)n * sizeof(uint64_t));
  hipLaunchKernelGGL(k_cancel_keys, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, p->stream,
                     d_bank, n, p->cm, p->ckeys.as<uint64_t>());
  HIPC(hipGetLastError());
  const int end_bit = bin_key_end_bit((uint64_t)p->cm.nx * (uint64_t)p->cm.ny * (uint64_t)p->cm.nz *
                                      (uint64_t)p->cm.ne);
  size_t temp = 0;
  HIPC(rocprim::radix_sort_keys(nullptr, temp, p->ckeys.as<uint64_t>(), p->ckeys2.as<uint64_t>(),
                                (size_t)n, 32, end_bit, p->stream));
  p->ctemp.reserve(std::max<size_t>(temp, 16));
  HIPC(rocprim::radix_sort_keys(p->ctemp.p, temp, p->ckeys.as<uint64_t>(), p->ckeys2.as<uint64_t>(),
                                (size_t)n, 32, end_bit, p->stream));
  const uint64_t* keys = p->ckeys2.as<uint64_t>();
  p->cw.reserve((size_t)n * 2 * sizeof(double));
  double* w = p->cw.as<double>();
  hipLaunchKernelGGL(k_cancel_gather, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, p->stream,
                     (const abx_banked*)d_bank, keys, n, w, w + n);
  HIPC(hipGetLastError());
  const int64_t chunks = (n + kCancelChunk - 1) / kCancelChunk;
  hipLaunchKernelGGL(k_cancel_bins, dim3((unsigned)std::min<int64_t>(chunks, 8192)), dim3(64), 0,
                     p->stream, keys, n, (const double*)w, (const double*)(w + n), d_bank);
  HIPC(hipGetLastError());
}

bool needs_secondaries(const HostTables& T) {
  if (T.st.tracking == ABX_CARTER_TRACKING) return true;
  if (T.st.mode != ABX_FIXED_SOURCE) return false;
  for (int f : T.fissile)
    if (f) return true;
  return false;
}

struct TransportOut {
  int64_t n_sites = 0;
  abx_scores scores{};
  abx_diag diag{};
};

// Device scalars of one transport call (scores, counters, error, staging
// count): a problem's own block, or a slice of a batch-wide block.
struct CallScalars {
  double* scores;               // [8]
  unsigned long long* counters; // [8]
  int32_t* err;
  unsigned long long* err_hist;
  unsigned long long* n_stage;
  unsigned long long* n_hole;
};

// Grows the problem's transport scratch for a call over n histories and
// fills the launch context (the scalars must be zeroed by the caller).
KCtx make_kctx(abx_problem* p, const abx_particle* d_in, int64_t n, double kcol, int converged,
               const CallScalars& cs, SecRec* sec, int sec_cap) {
  const HostTables& T = p->T;
  const bool banks = T.st.mode != ABX_FIXED_SOURCE;
  const int64_t stage_cap = banks ? 4 * n + 4096 : 0;
  if (banks) {
    p->stage.reserve((size_t)stage_cap * sizeof(abx_banked));
    p->stage_parent.reserve((size_t)stage_cap * sizeof(uint32_t));
    p->stage_daughter.reserve((size_t)stage_cap * sizeof(uint32_t));
    p->daughters.reserve((size_t)std::max<int64_t>(1, n) * sizeof(uint32_t));
  }
  KCtx K{};
  K.in = d_in;
  K.n = n;
  K.kcol = kcol;
  K.converged = converged;
  K.sec_cap = sec_cap;
  K.sec = sec_cap ? sec : nullptr;
  K.stage = banks ? p->stage.as<abx_banked>() : nullptr;
  K.stage_parent = banks ? p->stage_parent.as<uint32_t>() : nullptr;
  K.stage_daughter = banks ? p->stage_daughter.as<uint32_t>() : nullptr;
  K.n_stage = cs.n_stage;
  K.n_hole = cs.n_hole;
  K.stage_cap = stage_cap;
  K.daughters = banks ? p->daughters.as<uint32_t>() : nullptr;
  if (banks && p->exact) {
    p->stage_pinfo.reserve((size_t)stage_cap * sizeof(abx_parent_info));
    K.stage_pinfo = p->stage_pinfo.as<abx_parent_info>();
  }
  K.scores = cs.scores;
  K.counters = cs.counters;
  K.err = cs.err;
  K.err_hist = cs.err_hist;
  return K;
}

// Host copies of a call's scalars -> result; raises on a device error.
TransportOut collect_scalars(const double* sc, const unsigned long long* cn, int32_t err,
                             unsigned long long err_hist) {
  TransportOut res;
  res.scores = abx_scores{sc[0], sc[1], sc[2], sc[3], sc[4], sc[5]};
  res.diag.histories = cn[0];
  res.diag.collisions = cn[1];
  res.diag.virtual_collisions = cn[2];
  res.diag.flights = cn[3];
  res.diag.crossings = cn[4];
  res.diag.tally_updates = cn[5];
  res.diag.lost = cn[6];
  res.diag.secondaries = cn[7];
  if (err) {
    res.diag.error = err;
    res.diag.error_history = err_hist;
    fail(err, "transport failed at history " + std::to_string(err_hist));
  }
  return res;
}

// Orders the staged fission sites of a call over n parents into p->ordered
// (parent order, then push order); returns their number.  n_staged staging
// slots were handed out, n_holes of them hold no site.  With `pinfo` the
// sites' parent records go from p->stage_pinfo to p->ordered_pinfo alike.
int64_t order_fission_bank(abx_problem* p, int64_t n, uint64_t n_staged, uint64_t n_holes,
                           bool pinfo = false) {
  const int64_t stage_cap = 4 * n + 4096;
  if ((int64_t)n_staged > stage_cap) fail(ABX_E_CAPACITY, "fission staging overflow");
  p->offsets.reserve((size_t)std::max<int64_t>(1, n) * sizeof(uint64_t));
  const uint64_t total = device_scan(p, p->daughters.as<uint32_t>(), p->offsets.as<uint64_t>(), n);
  if (n_holes > n_staged || total != n_staged - n_holes)
    fail(ABX_E_INVALID, "fission site count mismatch");
  p->ordered.reserve((size_t)std::max<uint64_t>(1, total) * sizeof(abx_banked));
  if (pinfo) p->ordered_pinfo.reserve((size_t)std::max<uint64_t>(1, total) * sizeof(abx_parent_info));
  if (total) {
    p->err.reserve(sizeof(int32_t));
    HIPC(hipMemsetAsync(p->err.p, 0, sizeof(int32_t), p->stream));
    hipLaunchKernelGGL(k_order_bank, dim3((unsigned)blocks_for((int64_t)n_staged, 256)), dim3(256),
                       0, p->stream, p->stage.as<abx_banked>(), p->stage_parent.as<uint32_t>(),
                       p->stage_daughter.as<uint32_t>(), p->offsets.as<uint64_t>(),
                       (int64_t)n_staged, n, total, p->ordered.as<abx_banked>(),
                       p->err.as<int32_t>(),
                       pinfo ? (const abx_parent_info*)p->stage_pinfo.p : nullptr,
                       pinfo ? p->ordered_pinfo.as<abx_parent_info>() : nullptr);
    HIPC(hipGetLastError());
    int32_t err = 0;
    HIPC(hipMemcpyAsync(&err, p->err.p, sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    if (err) fail(err, "fission bank ordering: a staging slot names no parent of this call");
  }
  p->last_sites = (int64_t)total;
  return (int64_t)total;
}

void add_scores(abx_scores& a, const abx_scores& b) {
  a.k_col += b.k_col;
  a.k_abs += b.k_abs;
  a.k_trk += b.k_trk;
  a.k_tot += b.k_tot;
  a.leakage += b.leakage;
  a.mig += b.mig;
}

// One transport() over a device-resident bank; ordered fission sites are
// left in p->ordered.  Tallies accumulate into tally_gen.
TransportOut transport_device(abx_problem* p, const abx_particle* d_in, int64_t n, double kcol,
                              int converged, abx_trace_rec* d_trace, int trace_cap,
                              int32_t* d_nrec) {
  const HostTables& T = p->T;
  if (d_trace && T.st.precision == ABX_SPLIT)
    fail(ABX_E_INVALID, "event traces are recorded in FP64 precision only");
  if (p->exact && p->engine != ABX_ENGINE_HISTORY)
    fail(ABX_E_INVALID, "exact cancellation needs the history engine (parent records)");
  p->scores.reserve(8 * sizeof(double));
  p->counters.reserve(8 * sizeof(unsigned long long));
  p->err.reserve(sizeof(int32_t));
  p->err_hist.reserve(sizeof(unsigned long long));
  p->n_stage.reserve(2 * sizeof(unsigned long long));  // slots handed out, holes
  HIPC(hipMemsetAsync(p->scores.p, 0, 8 * sizeof(double), p->stream));
  HIPC(hipMemsetAsync(p->counters.p, 0, 8 * sizeof(unsigned long long), p->stream));
  HIPC(hipMemsetAsync(p->err.p, 0, sizeof(int32_t), p->stream));
  HIPC(hipMemsetAsync(p->n_stage.p, 0, 2 * sizeof(unsigned long long), p->stream));
  p->last_sites = 0;
  const bool banks = T.st.mode != ABX_FIXED_SOURCE;
  const int sec_cap = needs_secondaries(T) ? 16 : 0;
  // The stack is indexed by persistent lane (history engine) or by history
  // (event engine).  The persistent grid has at most ceil(n / bs) blocks of
  // bs <= 256 lanes, and any lane may take any history, so size it for
  // n + 256 slots.
  if (sec_cap) p->sec.reserve((size_t)(std::max<int64_t>(1, n) + 256) * sec_cap * sizeof(SecRec));
  const CallScalars cs{p->scores.as<double>(), p->counters.as<unsigned long long>(),
                       p->err.as<int32_t>(), p->err_hist.as<unsigned long long>(),
                       p->n_stage.as<unsigned long long>(),
                       p->n_stage.as<unsigned long long>() + 1};
  KCtx K = make_kctx(p, d_in, n, kcol, converged, cs, p->sec.as<SecRec>(), sec_cap);
  K.trace = d_trace;
  K.trace_cap = trace_cap;
  K.n_rec = d_nrec;
  if (n > 0) {
    if (p->timing && !p->ke0) {
      HIPC(hipEventCreate(&p->ke0));
      HIPC(hipEventCreate(&p->ke1));
    }
    if (p->engine == ABX_ENGINE_HISTORY) {
      if (p->timing) HIPC(hipEventRecord(p->ke0, p->stream));
      p->pool_next.reserve(sizeof(unsigned long long));
      HIPC(history_transport(p->P, p->T.extended, K, p->stream, sec_cap > 0,
                             p->pool_next.as<unsigned long long>(), p->tuning));
      if (p->timing) {
        HIPC(hipEventRecord(p->ke1, p->stream));
        HIPC(hipEventSynchronize(p->ke1));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, p->ke0, p->ke1));
        p->kernel_ms += ms;
        p->kernel_launches++;
        p->kernel_items += n;
      }
    } else {
      double ms = 0.0;
      int64_t launches = 0;
      HIPC(event_transport(p->P, K, p->ev, p->stream, sec_cap > 0, p->timing, &ms, &launches));
      p->kernel_ms += ms;
      p->kernel_launches += launches;
      p->kernel_items += p->ev.flight_items;
    }
  }
  // read back scalars
  double sc[8];
  unsigned long long cn[8];
  int32_t err = 0;
  unsigned long long err_hist = 0, nst[2] = {0, 0};
  HIPC(hipMemcpyAsync(sc, p->scores.p, 8 * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipMemcpyAsync(cn, p->counters.p, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipMemcpyAsync(&err, p->err.p, sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipMemcpyAsync(&err_hist, p->err_hist.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipMemcpyAsync(nst, p->n_stage.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                      p->stream));
  HIPC(hipStreamSynchronize(p->stream));
  TransportOut res = collect_scalars(sc, cn, err, err_hist);
  if (banks) res.n_sites = order_fission_bank(p, n, nst[0], nst[1], K.stage_pinfo != nullptr);
  add_scores(p->gen_scores, res.scores);
  return res;
}

void upd(int gen, double x, double& avg, double& var) {
  const double dgen = (double)gen;
  const double avg_old = avg, var_old = var;
  avg = avg_old + (x - avg_old) / dgen;
  if (gen > 1) var = var_old + ((x - avg_old) * (x - avg_old) / dgen) - (var_old / (dgen - 1.));
}

// The scalar part of a record (Tallies::record_generation,
// src/tallies.cpp:159-203) and the tally generation count.
void score_values(const abx_scores& s, double W, double x[6]) {
  x[0] = s.k_col / W;
  x[1] = s.k_abs / W;
  x[2] = s.k_trk / W;
  x[3] = s.k_tot / W;
  x[4] = s.leakage / W;
  x[5] = s.mig / W;
}

// Tallies::calc_gen_values: the generation's values join the per-generation
// vectors (every generation, recorded or not).
void push_gen_values(abx_problem* p, const abx_scores& s, double W) {
  double x[6];
  score_values(s, W, x);
  for (int q = 0; q < 6; q++) p->gen_vals[q].push_back(x[q]);
}

void record_scores(abx_problem* p, const abx_scores& s, double W) {
  p->gen++;
  double x[6];
  score_values(s, W, x);
  for (int q = 0; q < 6; q++) upd(p->gen, x[q], p->s_avg[q], p->s_var[q]);
  p->tally_g++;
}

// MeshTally::record_generation (src/mesh_tally.cpp:121-150) of every mesh
// tally, as record number p->tally_g.
void record_tallies(abx_problem* p, double multiplier, hipStream_t st = nullptr,
                    double* gen = nullptr) {
  const int64_t nb = p->T.tally_bins;
  if (nb > 0) {
    hipLaunchKernelGGL(k_record_generation, dim3((unsigned)blocks_for(nb, 256)), dim3(256), 0,
                       st ? st : p->stream, gen ? gen : p->tally_gen.as<double>(),
                       p->tally_avg.as<double>(), p->tally_var.as<double>(), nb,
                       (double)p->tally_g, 0, multiplier);
    HIPC(hipGetLastError());
  }
}

void record_generation(abx_problem* p, const abx_scores& s, double W,
                       hipStream_t st = nullptr, double* gen = nullptr) {
  record_scores(p, s, W);
  record_tallies(p, 1.0, st, gen);
}

void clear_generation(abx_problem* p, hipStream_t st = nullptr, double* gen = nullptr) {
  p->gen_scores = abx_scores{};
  const int64_t nb = p->T.tally_bins;
  if (nb > 0) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks_for(nb, 256)), dim3(256), 0,
                       st ? st : p->stream, gen ? gen : p->tally_gen.as<double>(), nb, 0.0);
    HIPC(hipGetLastError());
  }
}

void sample_sources_device(abx_problem* p, uint64_t first, int64_t n, abx_particle* d_out) {
  HIPC(hipMemsetAsync(p->err.p, 0, sizeof(int32_t), p->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_sample_sources, dim3((unsigned)blocks_for(n, 256)), dim3(256),
                       (size_t)256 * trk_bytes(p->P.trk_D, p->P.trk_L), p->stream,
                       p->P, p->S, first, n, d_out, p->err.as<int32_t>(),
                       p->err_hist.as<unsigned long long>());
    HIPC(hipGetLastError());
  }
  int32_t err = 0;
  unsigned long long h = 0;
  HIPC(hipMemcpyAsync(&err, p->err.p, sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipMemcpyAsync(&h, p->err_hist.p, sizeof(h), hipMemcpyDeviceToHost, p->stream));
  HIPC(hipStreamSynchronize(p->stream));
  if (err) fail(err, "source sampling failed at history " + std::to_string(h));
}

void add_diag(abx_diag& a, const abx_diag& b) {
  a.histories += b.histories;
  a.collisions += b.collisions;
  a.virtual_collisions += b.virtual_collisions;
  a.flights += b.flights;
  a.crossings += b.crossings;
  a.tally_updates += b.tally_updates;
  a.lost += b.lost;
  a.secondaries += b.secondaries;
}

}  // namespace

#define ABX_TRY(...)                           \
  try {                                        \
    __VA_ARGS__;                               \
    return ABX_OK;                             \
  } catch (const AbxError& e) {                \
    g_err = e.msg;                             \
    return e.code;                             \
  } catch (const TableError& e) {              \
    g_err = e.what();                          \
    return e.code;                             \
  } catch (const std::exception& e) {          \
    g_err = e.what();                          \
    return ABX_E_INVALID;                      \
  }

extern "C" {

const char* abx_last_error(void) { return g_err.c_str(); }
const char* abx_version(void) { return "abeille-amd 0.1 (gfx950)"; }

int abx_create(const abx_problem_desc* desc, abx_problem** out) {
  abx_problem* p = nullptr;
  try {
    if (!out) fail(ABX_E_INVALID, "null output handle");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) fail(ABX_E_NO_DEVICE, "no HIP device");
    p = new abx_problem;
    HIPC(hipGetDevice(&p->device));
    build_tables(desc, p->T);
    HIPC(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    const HostTables& T = p->T;
    p->tally_gen.reserve((size_t)std::max<int64_t>(1, T.tally_bins) * sizeof(double));
    p->tally_avg.reserve((size_t)std::max<int64_t>(1, T.tally_bins) * sizeof(double));
    p->tally_var.reserve((size_t)std::max<int64_t>(1, T.tally_bins) * sizeof(double));
    const size_t tb = (size_t)std::max<int64_t>(1, T.tally_bins) * sizeof(double);
    HIPC(hipMemset(p->tally_gen.p, 0, tb));
    HIPC(hipMemset(p->tally_avg.p, 0, tb));
    HIPC(hipMemset(p->tally_var.p, 0, tb));
    // all tables in one 16-byte-aligned block: pass 1 sizes it, pass 2
    // fills a host image that is uploaded in one copy
    size_t blob_bytes = 0;
    auto size_map = [&blob_bytes](const auto& v) {
      using E = typename std::decay_t<decltype(v)>::value_type;
      blob_bytes = (blob_bytes + 15) & ~(size_t)15;
      blob_bytes += std::max<size_t>(1, v.size()) * sizeof(E);
      return (const E*)nullptr;
    };
    (void)dev_view(T, size_map, nullptr);
    blob_bytes = (blob_bytes + 15) & ~(size_t)15;
    p->blob.reserve(std::max<size_t>(16, blob_bytes));
    std::vector<unsigned char> image(std::max<size_t>(16, blob_bytes), 0);
    size_t off = 0;
    unsigned char* dbase = p->blob.as<unsigned char>();
    auto fill_map = [&](const auto& v) {
      using E = typename std::decay_t<decltype(v)>::value_type;
      off = (off + 15) & ~(size_t)15;
      const E* d = (const E*)(dbase + off);
      if (!v.empty()) std::memcpy(image.data() + off, v.data(), v.size() * sizeof(E));
      off += std::max<size_t>(1, v.size()) * sizeof(E);
      return d;
    };
    p->P = dev_view(T, fill_map, p->tally_gen.as<double>());
    HIPC(hipMemcpy(p->blob.p, image.data(), image.size(), hipMemcpyHostToDevice));
    p->P.blob = dbase;
    p->P.blob_bytes = (int64_t)blob_bytes;
    std::vector<double> w, cdf;
    for (auto& s : T.sources) w.push_back(s.weight);
    cdf.resize(w.size());
    build_discrete_cdf(w.data(), (int)w.size(), cdf.data());
    p->S.src = p->upload(T.sources);
    p->S.cdf = p->upload(cdf);
    p->S.n = (int32_t)w.size();
    p->err.reserve(sizeof(int32_t));
    p->err_hist.reserve(sizeof(unsigned long long));
    if (desc->cancel) {
      const abx_cancel_mesh& c = *desc->cancel;
      if (!(c.low[0] < c.high[0] && c.low[1] < c.high[1] && c.low[2] < c.high[2]))
        fail(ABX_E_INVALID, "Low position is not lower than hi position in ApproximateMeshCancelator.");
      if (c.shape[0] < 1 || c.shape[1] < 1 || c.shape[2] < 1) fail(ABX_E_INVALID, "cancellation shape");
      if (c.n_ebounds == 1 || c.n_ebounds < 0) fail(ABX_E_INVALID, "energy_edges must have at least two entries");
      p->cancel = true;
      p->cm.lx = c.low[0];
      p->cm.ly = c.low[1];
      p->cm.lz = c.low[2];
      p->cm.nx = c.shape[0];
      p->cm.ny = c.shape[1];
      p->cm.nz = c.shape[2];
      p->cm.dx = (c.high[0] - c.low[0]) / (double)(uint32_t)c.shape[0];
      p->cm.dy = (c.high[1] - c.low[1]) / (double)(uint32_t)c.shape[1];
      p->cm.dz = (c.high[2] - c.low[2]) / (double)(uint32_t)c.shape[2];
      p->cm.n_edges = c.n_ebounds;
      p->cm.ne = c.n_ebounds > 0 ? c.n_ebounds - 1 : 1;
      std::vector<double> edges(c.ebounds, c.ebounds + std::max(0, c.n_ebounds));
      p->cm.edges = p->upload(edges);
      if ((int64_t)c.shape[0] * c.shape[1] * c.shape[2] * p->cm.ne >= 0xffffffffll)
        fail(ABX_E_INVALID, "cancellation mesh too large");
    }
    if (T.exact) {  // validated by build_tables
      p->exact = T.exact;
      DevExact& E = p->dex;
      E.kind = T.exact;
      E.beta = T.ex.beta;
      E.n_samples = T.ex.n_samples;
      E.chi_matrix = T.ex.chi_matrix != 0;
      E.use_virtual = T.ex.use_virtual_collisions != 0;
      E.sobol = T.ex.sobol != 0;
      for (int k = 0; k < 4; k++) E.S[k] = T.ex_shape[k];
      E.n_gbins = T.exact == ABX_EXACT_MG ? T.ex.n_group_bins : 0;
      E.nmat = std::max(1, T.nmat);
      E.G = T.G;
      for (int k = 0; k < 3; k++) {
        E.low[k] = T.ex.low[k];
        E.hi[k] = T.ex.high[k];
        E.pitch[k] = T.ex_pitch[k];
      }
      E.gb_off = p->upload(T.ex_gb_off);
      E.gb = p->upload(T.ex_gb);
      E.g2bin = p->upload(T.ex_g2bin);
      E.chi = p->upload(E.chi_matrix ? T.chi_norm : std::vector<double>{});
      std::vector<uint64_t> sm((size_t)kSobolDims * kSobolBits);
      sobol_build(sm.data());
      E.sobolM = p->upload(sm);
    }
    p->grng = pcg_seed(T.st.seed);  // settings::rng (src/settings.cpp:116-118)
    *out = p;
    return ABX_OK;
  } catch (const AbxError& e) {
    g_err = e.msg;
    delete p;
    return e.code;
  } catch (const TableError& e) {
    g_err = e.what();
    delete p;
    return e.code;
  }
}

void abx_destroy(abx_problem* p) { delete p; }

int abx_set_engine(abx_problem* p, int engine) {
  ABX_TRY({
    if (!p || (engine != ABX_ENGINE_EVENT && engine != ABX_ENGINE_HISTORY)) fail(ABX_E_INVALID, "engine");
    p->engine = engine;
  });
}

int abx_synchronize(void) {
  ABX_TRY({ HIPC(hipDeviceSynchronize()); });
}

int abx_set_tuning(abx_problem* p, const abx_tuning* t) {
  ABX_TRY({
    if (!p || !t) fail(ABX_E_INVALID, "null argument");
    if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256)
      fail(ABX_E_INVALID, "block_size must be 64, 128 or 256");
    if (t->blocks_per_cu < 0 || t->tail_threshold < 0 || t->sort_every < 0)
      fail(ABX_E_INVALID, "tuning values");
    p->tuning = *t;
    p->ev.tail_threshold = t->tail_threshold;
    p->ev.sort_every = t->sort_every;
  });
}

int abx_time_transport(abx_problem* p, size_t n, int converged, double* seconds) {
  ABX_TRY({
    if (!p || !seconds || n == 0) fail(ABX_E_INVALID, "timing arguments");
    // a scratch bank of fresh sources from history ids the runs do not use
    p->in.reserve(n * sizeof(abx_particle));
    sample_sources_device(p, (uint64_t)1 << 62, (int64_t)n, p->in.as<abx_particle>());
    const abx_scores saved = p->gen_scores;
    if (!p->ke0) {
      HIPC(hipEventCreate(&p->ke0));
      HIPC(hipEventCreate(&p->ke1));
    }
    const bool timing = p->timing;
    p->timing = false;
    HIPC(hipEventRecord(p->ke0, p->stream));
    (void)transport_device(p, p->in.as<abx_particle>(), (int64_t)n, p->run_kcol, converged, nullptr,
                           0, nullptr);
    HIPC(hipEventRecord(p->ke1, p->stream));
    HIPC(hipEventSynchronize(p->ke1));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, p->ke0, p->ke1));
    p->timing = timing;
    *seconds = ms * 1e-3;
    // leave the current generation's tallies and scores as they were
    clear_generation(p);
    p->gen_scores = saved;
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_get_tuning(abx_problem* p, abx_tuning* t) {
  ABX_TRY({
    if (!p || !t) fail(ABX_E_INVALID, "null argument");
    *t = p->tuning;
  });
}

int abx_sample_sources(abx_problem* p, uint64_t first_history, size_t n, abx_particle* out) {
  ABX_TRY({
    if (!p || (!out && n)) fail(ABX_E_INVALID, "null argument");
    p->in.reserve(std::max<size_t>(1, n) * sizeof(abx_particle));
    sample_sources_device(p, first_history, (int64_t)n, p->in.as<abx_particle>());
    HIPC(hipMemcpyAsync(out, p->in.p, n * sizeof(abx_particle), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
  });
}

static int transport_host(abx_problem* p, const abx_particle* bank, size_t n, double kcol,
                          int converged, abx_banked* fission_out, size_t cap, size_t* n_out,
                          abx_scores* scores, abx_diag* diag, abx_trace_rec* trace,
                          int cap_per_history, int32_t* n_rec) {
  ABX_TRY({
    if (!p || (!bank && n)) fail(ABX_E_INVALID, "null argument");
    if (n > 0xffffffffull) fail(ABX_E_INVALID, "bank larger than 2^32 histories");
    p->in.reserve(std::max<size_t>(1, n) * sizeof(abx_particle));
    if (n) HIPC(hipMemcpyAsync(p->in.p, bank, n * sizeof(abx_particle), hipMemcpyHostToDevice, p->stream));
    abx_trace_rec* d_trace = nullptr;
    int32_t* d_nrec = nullptr;
    if (trace) {
      if (cap_per_history <= 0 || !n_rec) fail(ABX_E_INVALID, "trace arguments");
      p->trace.reserve(std::max<size_t>(1, n * (size_t)cap_per_history) * sizeof(abx_trace_rec));
      p->n_rec.reserve(std::max<size_t>(1, n) * sizeof(int32_t));
      d_trace = p->trace.as<abx_trace_rec>();
      d_nrec = p->n_rec.as<int32_t>();
    }
    TransportOut r = transport_device(p, p->in.as<abx_particle>(), (int64_t)n, kcol, converged,
                                      d_trace, cap_per_history, d_nrec);
    if (n_out) *n_out = (size_t)r.n_sites;
    // the scores are the call's whatever happens to the sites below
    if (scores) *scores = r.scores;
    if (diag) *diag = r.diag;
    if (r.n_sites > 0) {
      // the ordered sites stay on the device: abx_fission_fetch can still
      // hand them over after this capacity error
      if (!fission_out || (size_t)r.n_sites > cap)
        fail(ABX_E_CAPACITY, "fission_out holds " + std::to_string(cap) + " sites, " +
                                 std::to_string(r.n_sites) + " were banked");
      HIPC(hipMemcpyAsync(fission_out, p->ordered.p, (size_t)r.n_sites * sizeof(abx_banked),
                          hipMemcpyDeviceToHost, p->stream));
    }
    if (trace) {
      HIPC(hipMemcpyAsync(trace, d_trace, n * (size_t)cap_per_history * sizeof(abx_trace_rec),
                          hipMemcpyDeviceToHost, p->stream));
      HIPC(hipMemcpyAsync(n_rec, d_nrec, n * sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
    }
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_transport(abx_problem* p, const abx_particle* bank, size_t n, double kcol, int converged,
                  abx_banked* fission_out, size_t cap, size_t* n_out, abx_scores* scores,
                  abx_diag* diag) {
  return transport_host(p, bank, n, kcol, converged, fission_out, cap, n_out, scores, diag,
                        nullptr, 0, nullptr);
}

int abx_transport_trace(abx_problem* p, const abx_particle* bank, size_t n, double kcol,
                        int converged, abx_banked* fission_out, size_t cap, size_t* n_out,
                        abx_scores* scores, abx_trace_rec* trace, int cap_per_history,
                        int32_t* n_rec) {
  if (!trace) {
    g_err = "trace buffer is required";
    return ABX_E_INVALID;
  }
  return transport_host(p, bank, n, kcol, converged, fission_out, cap, n_out, scores, nullptr,
                        trace, cap_per_history, n_rec);
}

int abx_fission_fetch(abx_problem* p, abx_banked* out, size_t cap, size_t* n_out) {
  ABX_TRY({
    if (!p) fail(ABX_E_INVALID, "null problem");
    const size_t n = (size_t)p->last_sites;
    if (n_out) *n_out = n;
    if (n > 0) {
      if (!out || n > cap)
        fail(ABX_E_CAPACITY, "fission_out holds " + std::to_string(cap) + " sites, " +
                                 std::to_string(n) + " were banked");
      HIPC(hipMemcpyAsync(out, p->ordered.p, n * sizeof(abx_banked), hipMemcpyDeviceToHost,
                          p->stream));
      HIPC(hipStreamSynchronize(p->stream));
    }
  });
}

int abx_tally_size(abx_problem* p, int tally, size_t* n_bins) {
  ABX_TRY({
    if (!p || !n_bins || tally < 0 || tally >= (int)p->T.tally.size()) fail(ABX_E_INVALID, "tally index");
    const DTally& t = p->T.tally[tally];
    *n_bins = (size_t)t.Ne * t.Nx * t.Ny * t.Nz;
  });
}

int abx_tally_fetch(abx_problem* p, int tally, double* gen_out) {
  ABX_TRY({
    size_t nb = 0;
    if (abx_tally_size(p, tally, &nb) != ABX_OK) fail(ABX_E_INVALID, "tally index");
    const DTally& t = p->T.tally[tally];
    HIPC(hipMemcpyAsync(gen_out, p->tally_gen.as<double>() + t.data_off, nb * sizeof(double),
                        hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_tally_record(abx_problem* p, double multiplier) {
  ABX_TRY({
    if (!p) fail(ABX_E_INVALID, "null problem");
    p->tally_g++;
    record_tallies(p, multiplier);
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_tally_clear(abx_problem* p) {
  ABX_TRY({
    if (!p) fail(ABX_E_INVALID, "null problem");
    clear_generation(p);
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_tally_stats(abx_problem* p, int tally, double* avg, double* std_err) {
  ABX_TRY({
    size_t nb = 0;
    if (abx_tally_size(p, tally, &nb) != ABX_OK) fail(ABX_E_INVALID, "tally index");
    const DTally& t = p->T.tally[tally];
    std::vector<double> var(nb);
    HIPC(hipMemcpyAsync(avg, p->tally_avg.as<double>() + t.data_off, nb * sizeof(double),
                        hipMemcpyDeviceToHost, p->stream));
    HIPC(hipMemcpyAsync(var.data(), p->tally_var.as<double>() + t.data_off, nb * sizeof(double),
                        hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    for (size_t i = 0; i < nb; i++) std_err[i] = std::sqrt(var[i] / (double)p->tally_g);
  });
}

static void run_reset(abx_problem* p) {
  const HostTables& T = p->T;
  p->run_t0 = std::chrono::steady_clock::now();
  p->run_gens = 0;
  p->gen = 0;
  for (int q = 0; q < 6; q++) {
    p->s_avg[q] = p->s_var[q] = 0.0;
    p->gen_vals[q].clear();
  }
  p->tally_g = 0;
  const size_t tb = (size_t)std::max<int64_t>(1, T.tally_bins) * sizeof(double);
  HIPC(hipMemsetAsync(p->tally_avg.p, 0, tb, p->stream));
  HIPC(hipMemsetAsync(p->tally_var.p, 0, tb, p->stream));
  clear_generation(p);
  p->run_started = false;
  p->run_hist = 0;
  p->run_kcol = 1.0;
  p->bank_n = 0;
  p->rep_chains = 0;
  p->grng = pcg_seed(T.st.seed);
}

// Whether a run stops at this generation boundary: a stop was requested
// (the reference's signal handler and premature_kill, src/main.cpp:47-75,
// src/power_iterator.cpp:412-417, 695-716), or the time left on the run's
// budget is under two mean generation times (PowerIterator::out_of_time,
// src/power_iterator.cpp:719-737).  `done` generations of the run have
// completed and `in_flight` more are queued on the device; the request is
// consumed.
static bool stop_now(abx_problem* p, int done, int in_flight) {
  if (p->stop_requested.exchange(0) != 0) return true;
  if (p->max_time <= 0.0 || done < 1) return false;
  const double used =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - p->run_t0).count();
  const double t_avg = used / (double)done;
  return p->max_time - used - (double)in_flight * t_avg < 2.0 * t_avg;
}

// Tallies' running means and standard errors (src/tallies.cpp:268-280).
static void fill_statistics(const abx_problem* p, abx_run_stats* stats) {
  const double dg = (double)std::max(1, p->gen);
  double err[6];
  for (int q = 0; q < 6; q++) err[q] = std::sqrt(p->s_var[q] / dg);
  stats->kcol_avg = p->s_avg[0];
  stats->kcol_err = err[0];
  stats->kabs_avg = p->s_avg[1];
  stats->kabs_err = err[1];
  stats->ktrk_avg = p->s_avg[2];
  stats->ktrk_err = err[2];
  stats->ktot_avg = p->s_avg[3];
  stats->ktot_err = err[3];
  stats->leak_avg = p->s_avg[4];
  stats->leak_err = err[4];
  stats->mig_avg = p->s_avg[5];
  stats->mig_err = err[5];
  stats->active_generations = p->gen;
}

// Grows a buffer keeping its first `keep` bytes.
static void grow_keep(DBuf& b, size_t bytes, size_t keep, hipStream_t s) {
  if (bytes <= b.bytes) return;
  DBuf n;
  n.reserve(bytes);
  if (keep) HIPC(hipMemcpyAsync(n.p, b.p, keep, hipMemcpyDeviceToDevice, s));
  HIPC(hipStreamSynchronize(s));
  b.release();
  b = n;
}

// Exact / basic-exact cancellation (perform_regional_cancellation,
// src/power_iterator.cpp:751-777) of the n ordered fission sites in
// p->ordered with their parent records in p->ordered_pinfo: weights change
// in place and the uniform particles are appended; returns the new bank
// size.  settings::rng (p->grng) advances as the reference's does.
static int64_t exact_cancel_device(abx_problem* p, int64_t n) {
  const DevExact& E = p->dex;
  if (E.kind == ABX_BASIC_EXACT_MG && E.beta == ABX_BETA_ZERO) return n;
  if (n <= 0) return n;
  if (n > 0xffffffffll) fail(ABX_E_INVALID, "exact cancellation: more than 2^32 sites");
  hipStream_t s = p->stream;
  const size_t lds = (size_t)256 * trk_bytes(p->P.trk_D, p->P.trk_L);
  abx_banked* bank = p->ordered.as<abx_banked>();
  const abx_parent_info* par = p->ordered_pinfo.as<abx_parent_info>();
  require_records(p->ordered, n, sizeof(abx_banked), "cancelled bank");
  require_records(p->ordered_pinfo, n, sizeof(abx_parent_info), "parent records");
  p->ckeys.reserve((size_t)n * sizeof(uint64_t));
  p->ckeys2.reserve((size_t)n * sizeof(uint64_t));
  p->ex_head.reserve((size_t)n * sizeof(uint32_t));
  p->ex_offs.reserve((size_t)n * sizeof(uint64_t));
  HIPC(hipMemsetAsync(p->err.p, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_ex_keys, dim3((unsigned)blocks_for(n, 256)), dim3(256), lds, s, p->P, E,
                     (const abx_banked*)bank, par, n, p->ckeys.as<uint64_t>(), p->err.as<int32_t>());
  HIPC(hipGetLastError());
  const int end_bit = bin_key_end_bit((uint64_t)(E.S[0] + 1) * (uint64_t)(E.S[1] + 1) *
                                      (uint64_t)(E.S[2] + 1) * (uint64_t)(E.S[3] + 1) *
                                      (uint64_t)E.nmat);
  size_t temp = 0;
  HIPC(rocprim::radix_sort_keys(nullptr, temp, p->ckeys.as<uint64_t>(), p->ckeys2.as<uint64_t>(),
                                (size_t)n, 32, end_bit, s));
  p->ctemp.reserve(std::max<size_t>(temp, 16));
  HIPC(rocprim::radix_sort_keys(p->ctemp.p, temp, p->ckeys.as<uint64_t>(), p->ckeys2.as<uint64_t>(),
                                (size_t)n, 32, end_bit, s));
  const uint64_t* keys = p->ckeys2.as<uint64_t>();
  hipLaunchKernelGGL(k_ex_heads, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, s, keys, n,
                     p->ex_head.as<uint32_t>());
  HIPC(hipGetLastError());
  const int64_t nb = (int64_t)device_scan(p, p->ex_head.as<uint32_t>(), p->ex_offs.as<uint64_t>(), n);
  int32_t err = 0;
  HIPC(hipMemcpyAsync(&err, p->err.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (err) fail(err, "exact cancellation: a fission site lies in no cell");
  if (nb == 0) return n;
  p->ex_bins.reserve((size_t)nb * sizeof(ExBinD));
  ExBinD* bins = p->ex_bins.as<ExBinD>();
  const uint32_t* head = p->ex_head.as<uint32_t>();
  const uint64_t* offs = p->ex_offs.as<uint64_t>();
  hipLaunchKernelGGL(k_ex_bins, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, s, keys, n,
                     head, offs, bins);
  HIPC(hipGetLastError());
  // the per-site arrays, in sorted order
  p->ex_fv.reserve((size_t)n * 11 * sizeof(double));
  ExSiteArrays A;
  double* a0 = p->ex_fv.as<double>();
  double** fields[] = {&A.w, &A.w2, &A.fv, &A.fi, &A.cc, &A.cw, &A.cw2, &A.kp1, &A.up1, &A.kp2, &A.up2};
  for (int k = 0; k < 11; k++) *fields[k] = a0 + (size_t)k * (size_t)n;
  // site kernels: 256 lanes a block; region kernels: one 64-lane block a region
  const dim3 gs((unsigned)blocks_for(n, 256)), b256(256), gr((unsigned)nb), b64(64);
  hipLaunchKernelGGL(k_ex_site, gs, b256, 0, s, (const abx_banked*)bank, keys, n, head, offs, bins,
                     A);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_ex_region, gr, b64, 0, s, E, bins, nb, A);
  HIPC(hipGetLastError());
  // basic: the per-region streams skip n_samples * 100 (x2 for average-g)
  // draws per site of every earlier region (:475-489)
  const bool basic = E.kind == ABX_BASIC_EXACT_MG;
  const uint64_t max_rn = (uint64_t)E.n_samples * 100u * (E.beta == ABX_BETA_AVERAGE_G ? 2u : 1u);
  p->ex_smp.reserve((size_t)nb * (size_t)E.n_samples * sizeof(ExSample));
  hipLaunchKernelGGL(k_ex_samples, dim3((unsigned)blocks_for(nb, 256)), dim3(256), lds, s, p->P, E,
                     bins, nb, p->grng, max_rn, p->ex_smp.as<ExSample>());
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_ex_f, gs, b256, 0, s, p->P, E, par, keys, n, head, offs, bins,
                     (const ExSample*)p->ex_smp.p, A);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_ex_sums, gr, b64, 0, s, E, bins, nb, A);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_ex_p, gs, b256, 0, s, p->P, E, (const abx_banked*)bank, par, keys, n, head,
                     offs, bins, A);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_ex_uniform, gr, b64, 0, s, bins, nb, A);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_ex_apply, gs, b256, 0, s, bank, keys, n, head, offs, (const ExBinD*)bins, A);
  HIPC(hipGetLastError());
  if (basic && E.beta != ABX_BETA_MINIMUM) {
    ExBinD last{};
    HIPC(hipMemcpyAsync(&last, bins + (nb - 1), sizeof(ExBinD), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    p->grng = pcg_advance(p->grng, (uint64_t)last.end * max_rn);
  }
  // get_new_particles
  p->ex_cnt.reserve((size_t)nb * sizeof(uint32_t));
  hipLaunchKernelGGL(k_ex_counts, dim3((unsigned)blocks_for(nb, 256)), dim3(256), 0, s,
                     (const ExBinD*)bins, nb, p->ex_cnt.as<uint32_t>());
  HIPC(hipGetLastError());
  const uint64_t nnew = device_scan(p, p->ex_cnt.as<uint32_t>(), p->ex_offs.as<uint64_t>(), nb);
  if (nnew == 0) return n;
  grow_keep(p->ordered, (size_t)(n + (int64_t)nnew) * sizeof(abx_banked), (size_t)n * sizeof(abx_banked), s);
  p->ex_state.reserve(sizeof(uint64_t));
  HIPC(hipMemsetAsync(p->err.p, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_ex_new, dim3(1), dim3(256), lds, s, p->P, E, (const ExBinD*)bins, nb,
                     (const uint64_t*)p->ex_offs.p, p->grng, p->ordered.as<abx_banked>() + n,
                     p->ex_state.as<uint64_t>(), p->err.as<int32_t>());
  HIPC(hipGetLastError());
  uint64_t g = 0;
  HIPC(hipMemcpyAsync(&g, p->ex_state.p, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(&err, p->err.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (err) fail(err, "Couldn't sample position for uniform particle.");
  p->grng = g;
  return n + (int64_t)nnew;
}

// PowerIterator::run with every generation a replay of captured graphs:
// graph 1 = zero the generation's scalars + the persistent transport kernel
// (bank and k read from GenDev), graph 2 = device scan of the daughter
// counts, ordered fission bank, weight sums, statistics, re-bank into the
// other buffer, tally statistics.  No host round trip inside a generation;
// HIP events between the two graphs time the transport kernel.  Used for
// the history engine without cancellation or traces; returns false
// otherwise.
static bool run_generations_graph(abx_problem* p, int ngenerations, int nignored,
                                  double* kcol_hist, abx_run_stats* stats) {
  const HostTables& T = p->T;
  if (!p->use_graphs || T.st.mode != ABX_K_EIGENVALUE || p->engine != ABX_ENGINE_HISTORY ||
      p->cancel || p->exact || !p->run_started)
    return false;
  const int64_t N = T.st.nparticles;
  const double W = (double)N;
  const uint64_t cap = (uint64_t)(4 * N + 4096);
  if (p->bank_n > cap) return false;
  hipStream_t s = p->stream;
  // buffers at capacity (the live bank is kept)
  grow_keep(p->bank_a, cap * sizeof(abx_particle), p->bank_n * sizeof(abx_particle), s);
  p->bank_b.reserve(cap * sizeof(abx_particle));
  const int64_t nb = (int64_t)((cap + 1023) / 1024);
  p->offsets.reserve(cap * sizeof(uint64_t));
  p->block_sums.reserve((size_t)(nb + 1) * sizeof(uint64_t));
  p->ordered.reserve(cap * sizeof(abx_banked));
  p->gdev.reserve(sizeof(GenDev));
  p->khist_dev.reserve((size_t)6 * std::max(1, ngenerations) * sizeof(double));
  p->pool_next.reserve(sizeof(unsigned long long));
  GenDev* g = p->gdev.as<GenDev>();
  GenDev h{};
  h.n = p->bank_n;
  h.run_hist = p->run_hist;
  h.cap = cap;
  h.kcol = p->run_kcol;
  h.gen = p->gen;
  for (int q = 0; q < 6; q++) {
    h.s_avg[q] = p->s_avg[q];
    h.s_var[q] = p->s_var[q];
  }
  h.hist_cap = std::max(1, ngenerations);
  HIPC(hipMemcpyAsync(g, &h, sizeof(GenDev), hipMemcpyHostToDevice, s));
  HIPC(hipStreamSynchronize(s));
  // transport context: counts and scalars live in GenDev
  const CallScalars cs{g->scores, g->counters, &g->err, &g->err_hist, &g->n_stage, &g->n_hole};
  KCtx K = make_kctx(p, p->bank_a.as<abx_particle>(), N, 1.0, 1, cs, nullptr, 0);
  p->daughters.reserve(cap * sizeof(uint32_t));
  K.daughters = p->daughters.as<uint32_t>();
  K.stage_cap = (int64_t)cap;
  K.n_dev = &g->n;
  K.kcol_dev = &g->kcol;
  const abx_tuning tu = p->tuning;
  const int bs = tu.block_size;
  const size_t cur = cursor_lds_bytes(bs, p->P.trk_D, p->P.trk_L);
  const int trk = T.st.tracking;
  const bool split = T.st.precision == ABX_SPLIT;
  const bool sec = needs_secondaries(T);
  if (sec) return false;  // k-eigenvalue without carter tracking only
  bool stage = false;
  int64_t grid = 0;
  // (carter tracking never gets here: it carries the secondary stack)
  const bool ext = p->T.extended;
  HIPC(dispatch_variant(trk, false, false, split, ext, [&](auto TT, auto, auto SE, auto SP, auto EX) {
    constexpr int TK = decltype(TT)::value;
    constexpr bool SEC = decltype(SE)::value, SPL = decltype(SP)::value, EXT = decltype(EX)::value;
    stage = tu.stage_tables && stage_pays(k_history<TK, false, SEC, SPL, true, false, EXT>,
                                          k_history<TK, false, SEC, SPL, false, false, EXT>, bs,
                                          cur, (size_t)p->P.blob_bytes);
    grid = stage ? resident_blocks(k_history<TK, false, SEC, SPL, true, false, EXT>, bs,
                                   cur + (size_t)p->P.blob_bytes)
                 : resident_blocks(k_history<TK, false, SEC, SPL, false, false, EXT>, bs, cur);
    return hipSuccess;
  }));
  const size_t lds = stage ? cur + (size_t)p->P.blob_bytes : cur;
  if (tu.blocks_per_cu > 0) {
    int dev = 0, cus = 0;
    HIPC(hipGetDevice(&dev));
    HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = std::min<int64_t>(grid, (int64_t)cus * tu.blocks_per_cu);
  }
  grid = std::min<int64_t>(grid, (int64_t)((cap + bs - 1) / bs));
  const int64_t tb = T.tally_bins;
  // capture: [parity][converged] x {transport, bookkeeping}
  hipGraphExec_t ex[2][2][2] = {};
  hipGraph_t gr[2][2][2] = {};
  auto capture = [&](int par, int conv) {
    abx_particle* in = par ? p->bank_b.as<abx_particle>() : p->bank_a.as<abx_particle>();
    abx_particle* out = par ? p->bank_a.as<abx_particle>() : p->bank_b.as<abx_particle>();
    KCtx Kc = K;
    Kc.in = in;
    Kc.converged = conv;
    HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_gen_begin, dim3(1), dim3(64), 0, s, g);
    HIPC(dispatch_variant(trk, false, false, split, ext,
                          [&](auto TT, auto, auto SE, auto SP, auto EX) {
      constexpr int TK = decltype(TT)::value;
      constexpr bool SEC = decltype(SE)::value, SPL = decltype(SP)::value;
      constexpr bool EXT = decltype(EX)::value;
      if (stage)
        hipLaunchKernelGGL((k_history<TK, false, SEC, SPL, true, false, EXT>),
                           dim3((unsigned)grid), dim3(bs), lds, s, p->P, Kc, &g->next);
      else
        hipLaunchKernelGGL((k_history<TK, false, SEC, SPL, false, false, EXT>),
                           dim3((unsigned)grid), dim3(bs), lds, s, p->P, Kc, &g->next);
      return hipSuccess;
    }));
    HIPC(hipStreamEndCapture(s, &gr[par][conv][0]));
    HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_scan_blocks_dev, dim3((unsigned)nb), dim3(1024), 0, s,
                       p->daughters.as<uint32_t>(), p->offsets.as<uint64_t>(),
                       p->block_sums.as<uint64_t>(), &g->n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, p->block_sums.as<uint64_t>(), nb,
                       &g->n_sites);
    hipLaunchKernelGGL(k_add_block_offsets_dev, dim3((unsigned)nb), dim3(1024), 0, s,
                       p->offsets.as<uint64_t>(), p->block_sums.as<uint64_t>(), &g->n);
    // grid-stride bookkeeping kernels: enough blocks to fill the GPU, not the capacity
    const unsigned gcap = (unsigned)std::min<uint64_t>((cap + 255) / 256, 2048);
    hipLaunchKernelGGL(k_order_bank_dev, dim3(gcap), dim3(256), 0, s, p->stage.as<abx_banked>(),
                       p->stage_parent.as<uint32_t>(), p->stage_daughter.as<uint32_t>(),
                       p->offsets.as<uint64_t>(), g, p->ordered.as<abx_banked>());
    hipLaunchKernelGGL(k_gen_end, dim3(1), dim3(64), 0, s, g, W, conv, p->khist_dev.as<double>());
    hipLaunchKernelGGL(k_rebank_dev, dim3(gcap), dim3(256), 0, s, p->ordered.as<abx_banked>(), g,
                       T.st.seed, T.st.stride, out);
    if (conv && tb > 0)
      hipLaunchKernelGGL(k_record_generation_dev, dim3((unsigned)blocks_for(tb, 256)), dim3(256), 0, s,
                         p->tally_gen.as<double>(), p->tally_avg.as<double>(),
                         p->tally_var.as<double>(), tb, g, 0);
    HIPC(hipStreamEndCapture(s, &gr[par][conv][1]));
    for (int k = 0; k < 2; k++)
      HIPC(hipGraphInstantiate(&ex[par][conv][k], gr[par][conv][k], nullptr, nullptr, 0));
  };
  // (destroys each graph once: an error after the first call lands in the
  // catch below, which calls it again)
  auto cleanup = [&]() {
    for (auto& a : ex)
      for (auto& b : a)
        for (auto& c : b)
          if (c) {
            (void)hipGraphExecDestroy(c);
            c = nullptr;
          }
    for (auto& a : gr)
      for (auto& b : a)
        for (auto& c : b)
          if (c) {
            (void)hipGraphDestroy(c);
            c = nullptr;
          }
  };
  try {
    while ((int)p->gen_events.size() < 2 * ngenerations + 2) {
      hipEvent_t e;
      HIPC(hipEventCreate(&e));
      p->gen_events.push_back(e);
    }
    bool converged = nignored == 0;
    int par = 0;
    int queued = ngenerations;  // generations launched (fewer after a stop)
    for (int gi = 0; gi < ngenerations; gi++) {
      if (gi >= 2) {
        // two generations ahead: gi - 2 has ended once gi - 1 started
        HIPC(hipEventSynchronize(p->gen_events[2 * (gi - 1)]));
        if (stop_now(p, p->run_gens + gi - 1, 1)) {
          queued = gi;
          break;
        }
      }
      const int conv = converged ? 1 : 0;
      if (!ex[par][conv][0]) capture(par, conv);
      HIPC(hipEventRecord(p->gen_events[2 * gi], s));
      HIPC(hipGraphLaunch(ex[par][conv][0], s));
      HIPC(hipEventRecord(p->gen_events[2 * gi + 1], s));
      HIPC(hipGraphLaunch(ex[par][conv][1], s));
      par ^= 1;
      if (gi + 1 == nignored) converged = true;
    }
    HIPC(hipEventRecord(p->gen_events[2 * queued], s));
    HIPC(hipMemcpyAsync(&h, g, sizeof(GenDev), hipMemcpyDeviceToHost, s));
    std::vector<double> hist((size_t)6 * ngenerations);
    HIPC(hipMemcpyAsync(hist.data(), p->khist_dev.p, hist.size() * sizeof(double),
                        hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    // the latest bank is in bank_b after an odd number of generations
    if (par) std::swap(p->bank_a, p->bank_b);
    float total = 0.f;
    double kms = 0.0;
    for (int gi = 0; gi < queued; gi++) {
      float x = 0.f;
      HIPC(hipEventElapsedTime(&x, p->gen_events[2 * gi], p->gen_events[2 * gi + 1]));
      kms += x;
    }
    HIPC(hipEventElapsedTime(&total, p->gen_events[0], p->gen_events[2 * queued]));
    cleanup();
    if (h.err) {
      if (h.err == -8) fail(ABX_E_INVALID, "No fission neutrons were produced.");
      fail(h.err, "transport failed at history " + std::to_string(h.err_hist));
    }
    p->gen = h.gen;
    p->tally_g = h.gen;
    for (int q = 0; q < 6; q++) {
      p->s_avg[q] = h.s_avg[q];
      p->s_var[q] = h.s_var[q];
    }
    // the generations that ran (a stop or a device error ends the run early)
    const int ran = std::min(queued, (int)h.hist_idx);
    for (int q = 0; q < 6; q++)
      p->gen_vals[q].insert(p->gen_vals[q].end(), hist.begin() + (size_t)q * ngenerations,
                            hist.begin() + (size_t)q * ngenerations + ran);
    if (kcol_hist) std::copy(hist.begin(), hist.begin() + ran, kcol_hist);
    p->run_gens += ran;
    p->last_sites = (int64_t)h.n_sites;
    p->bank_n = (size_t)h.n;
    p->run_hist = h.run_hist;
    p->run_kcol = h.kcol;
    p->gen_scores = abx_scores{};
    *stats = abx_run_stats{};
    fill_statistics(p, stats);
    abx_diag& d = stats->diag;
    d.histories = h.diag[0];
    d.collisions = h.diag[1];
    d.virtual_collisions = h.diag[2];
    d.flights = h.diag[3];
    d.crossings = h.diag[4];
    d.tally_updates = h.diag[5];
    d.lost = h.diag[6];
    d.secondaries = h.diag[7];
    stats->transport_seconds = total * 1e-3;
    stats->kernel_seconds = kms * 1e-3;
    stats->kernel_launches = queued;
    stats->generations_run = ran;
    stats->kernel_items = (int64_t)(d.flights + d.collisions);
    stats->histories = (int64_t)d.histories;
  } catch (...) {
    cleanup();
    throw;
  }
  return true;
}

// PowerIterator::run (src/power_iterator.cpp:305-429) for `R` independent
// chains of one k-eigenvalue problem that differ only in their seed, in one
// bank: every generation is one transport launch over all chains'
// histories plus the chain-aware bookkeeping, captured as two graphs and
// replayed.  Chain r samples histories 0..N-1 with seeds[r] and then runs
// exactly the histories abx_run with that seed runs; the mesh tallies and
// the running statistics follow the chains' mean per generation.
// The replica-mode kernel variants: surface or delta tracking, FP64 or
// split precision (no trace, no secondary stack).
extern "C++" {
template <bool EXT, class Fn>
static hipError_t rep_dispatch_t(int trk, bool split, Fn&& fn) {
  const BoolC<EXT> hx{};
  if (trk == 0)
    return split ? fn(IntC<0>{}, BoolC<true>{}, hx) : fn(IntC<0>{}, BoolC<false>{}, hx);
  if (trk == 1)
    return split ? fn(IntC<1>{}, BoolC<true>{}, hx) : fn(IntC<1>{}, BoolC<false>{}, hx);
  return hipErrorInvalidValue;
}
// fn(TRK, SPLIT, EXT) over the replica variants
template <class Fn>
static hipError_t rep_dispatch(int trk, bool split, bool ext, Fn&& fn) {
  return ext ? rep_dispatch_t<true>(trk, split, fn) : rep_dispatch_t<false>(trk, split, fn);
}
}

static void run_replicas(abx_problem* p, int R, const uint64_t* seeds, bool from_bank,
                         int ngenerations, int nignored, double* kcol_hist, abx_run_stats* stats,
                         bool cont) {
  const HostTables& T = p->T;
  if (T.st.mode != ABX_K_EIGENVALUE) fail(ABX_E_INVALID, "replicas are k-eigenvalue chains");
  if (p->cancel) fail(ABX_E_INVALID, "replicas: approximate cancellation is not supported");
  if (p->exact) fail(ABX_E_INVALID, "replicas: exact cancellation is not supported");
  if (needs_secondaries(T)) fail(ABX_E_INVALID, "replicas: carter tracking is not supported");
  // the bank every chain starts from (from_bank): the problem's current one
  const uint64_t n0 = from_bank ? (uint64_t)p->bank_n : 0;
  const double kcol0 = p->run_kcol;
  if (from_bank && (!p->run_started || n0 == 0))
    fail(ABX_E_INVALID, "replicas: no k-eigenvalue bank to start from (abx_run first)");
  if (cont) {
    if (p->rep_chains < 1) fail(ABX_E_INVALID, "replicas: no replica run to continue");
    R = p->rep_chains;
  } else {
    run_reset(p);
  }
  const int64_t N = T.st.nparticles;
  if (N < 1) fail(ABX_E_INVALID, "nparticles must be positive");
  const double W = (double)N;
  const uint64_t NR = (uint64_t)std::max<uint64_t>((uint64_t)N, n0) * (uint64_t)R;
  const uint64_t cap = 4 * NR + 4096;
  hipStream_t s = p->stream;
  if (from_bank) grow_keep(p->bank_a, cap * sizeof(abx_particle), n0 * sizeof(abx_particle), s);
  p->bank_a.reserve(cap * sizeof(abx_particle));
  p->bank_b.reserve(cap * sizeof(abx_particle));
  const int64_t nb = (int64_t)((cap + 1023) / 1024);
  p->offsets.reserve(cap * sizeof(uint64_t));
  p->block_sums.reserve((size_t)(nb + 1) * sizeof(uint64_t));
  p->ordered.reserve(cap * sizeof(abx_banked));
  p->gdev.reserve(sizeof(GenRep));
  p->khist_dev.reserve((size_t)(6 + R) * ngenerations * sizeof(double));
  GenRep* g = p->gdev.as<GenRep>();
  GenRep h{};
  if (cont) {
    // the chains' banks, k, history counters and statistics carry over;
    // the per-generation arrays and diagnostics are this call's
    HIPC(hipMemcpyAsync(&h, g, sizeof(GenRep), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (h.nrep != R) fail(ABX_E_INVALID, "replicas: device state does not match the run");
    h.hist_cap = ngenerations;
    h.hist_idx = 0;
    for (int k = 0; k < 8; k++) h.diag[k] = 0;
    HIPC(hipMemcpyAsync(g, &h, sizeof(GenRep), hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
  }
  uint64_t sd[kMaxReplicas];
  for (int r = 0; r < R; r++) sd[r] = seeds ? seeds[r] : T.st.seed + (uint64_t)r;
  // each chain's first generation: histories 0..N-1 sampled with its seed
  if (!cont) HIPC(hipMemsetAsync(p->err.p, 0, sizeof(int32_t), s));
  // from_bank: chain r's copy of the bank becomes its histories 0..n0-1 on
  // its own stream (the last chain first: chain 0's copy is the bank itself)
  for (int r = (cont || !from_bank) ? -1 : R - 1; r >= 0; r--) {
    hipLaunchKernelGGL(k_fork_chain, dim3((unsigned)blocks_for((int64_t)n0, 256)), dim3(256), 0,
                       s, p->bank_a.as<abx_particle>(), (int64_t)n0, sd[r], T.st.stride,
                       p->bank_a.as<abx_particle>() + (size_t)r * n0);
    HIPC(hipGetLastError());
  }
  for (int r = 0; r < (cont || from_bank ? 0 : R); r++) {
    DevProblem Pr = p->P;
    Pr.seed = sd[r];
    hipLaunchKernelGGL(k_sample_sources, dim3((unsigned)blocks_for(N, 256)), dim3(256),
                       (size_t)256 * trk_bytes(p->P.trk_D, p->P.trk_L), s, Pr, p->S, 0, N,
                       p->bank_a.as<abx_particle>() + (size_t)r * N, p->err.as<int32_t>(),
                       p->err_hist.as<unsigned long long>());
    HIPC(hipGetLastError());
  }
  if (!cont) {
    int32_t err = 0;
    HIPC(hipMemcpyAsync(&err, p->err.p, sizeof(err), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (err) fail(err, "source sampling failed");
    const uint64_t per = from_bank ? n0 : (uint64_t)N;  // histories per chain
    h = GenRep{};
    h.n = per * (uint64_t)R;
    h.cap = cap;
    h.nrep = R;
    for (int r = 0; r < R; r++) {
      h.seg[r] = (uint64_t)r * per;
      h.run_hist[r] = per;
      h.seed[r] = sd[r];
      h.kcol[r] = from_bank ? kcol0 : 1.0;
    }
    h.seg[R] = per * (uint64_t)R;
    h.hist_cap = ngenerations;
    HIPC(hipMemcpyAsync(g, &h, sizeof(GenRep), hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
  }
  const CallScalars cs{&g->scores[0][0], g->counters, &g->err, &g->err_hist, &g->n_stage,
                       &g->n_hole};
  KCtx K = make_kctx(p, p->bank_a.as<abx_particle>(), (int64_t)NR, 1.0, 1, cs, nullptr, 0);
  p->daughters.reserve(cap * sizeof(uint32_t));
  K.daughters = p->daughters.as<uint32_t>();
  K.stage_cap = (int64_t)cap;
  K.n_dev = &g->n;
  K.kcol_dev = nullptr;
  K.rep_start = g->seg;
  K.rep_kcol = g->kcol;
  K.rep_scores = &g->scores[0][0];
  K.nrep = R;
  const abx_tuning tu = p->tuning;
  const int bs = tu.block_size;
  const size_t cur = cursor_lds_bytes(bs, p->P.trk_D, p->P.trk_L);
  const int trk = T.st.tracking;
  const bool split = T.st.precision == ABX_SPLIT;
  bool stage = false;
  int64_t grid = 0;
  const bool ext = p->T.extended;
  HIPC(rep_dispatch(trk, split, ext, [&](auto TT, auto SP, auto EX) {
    constexpr int TK = decltype(TT)::value;
    constexpr bool SPL = decltype(SP)::value, EXT = decltype(EX)::value;
    stage = tu.stage_tables && stage_pays(k_history<TK, false, false, SPL, true, true, EXT>,
                                          k_history<TK, false, false, SPL, false, true, EXT>, bs,
                                          cur, (size_t)p->P.blob_bytes);
    grid = stage ? resident_blocks(k_history<TK, false, false, SPL, true, true, EXT>, bs,
                                   cur + (size_t)p->P.blob_bytes)
                 : resident_blocks(k_history<TK, false, false, SPL, false, true, EXT>, bs, cur);
    return hipSuccess;
  }));
  const size_t lds = stage ? cur + (size_t)p->P.blob_bytes : cur;
  if (tu.blocks_per_cu > 0) {
    int dev = 0, cus = 0;
    HIPC(hipGetDevice(&dev));
    HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = std::min<int64_t>(grid, (int64_t)cus * tu.blocks_per_cu);
  }
  grid = std::min<int64_t>(grid, (int64_t)((cap + bs - 1) / bs));
  const int64_t tb = T.tally_bins;
  double* hist_dev = p->khist_dev.as<double>();
  double* khist_dev = hist_dev + (size_t)6 * ngenerations;
  hipGraphExec_t ex[2][2][2] = {};
  hipGraph_t gr[2][2][2] = {};
  auto capture = [&](int par, int conv) {
    abx_particle* in = par ? p->bank_b.as<abx_particle>() : p->bank_a.as<abx_particle>();
    abx_particle* out = par ? p->bank_a.as<abx_particle>() : p->bank_b.as<abx_particle>();
    KCtx Kc = K;
    Kc.in = in;
    Kc.converged = conv;
    HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_rep_begin, dim3(1), dim3(256), 0, s, g);
    HIPC(rep_dispatch(trk, split, ext, [&](auto TT, auto SP, auto EX) {
      constexpr int TK = decltype(TT)::value;
      constexpr bool SPL = decltype(SP)::value, EXT = decltype(EX)::value;
      if (stage)
        hipLaunchKernelGGL((k_history<TK, false, false, SPL, true, true, EXT>),
                           dim3((unsigned)grid), dim3(bs), lds, s, p->P, Kc, &g->next);
      else
        hipLaunchKernelGGL((k_history<TK, false, false, SPL, false, true, EXT>),
                           dim3((unsigned)grid), dim3(bs), lds, s, p->P, Kc, &g->next);
      return hipSuccess;
    }));
    HIPC(hipStreamEndCapture(s, &gr[par][conv][0]));
    HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_scan_blocks_dev, dim3((unsigned)nb), dim3(1024), 0, s,
                       p->daughters.as<uint32_t>(), p->offsets.as<uint64_t>(),
                       p->block_sums.as<uint64_t>(), &g->n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, p->block_sums.as<uint64_t>(), nb,
                       &g->n_sites);
    hipLaunchKernelGGL(k_add_block_offsets_dev, dim3((unsigned)nb), dim3(1024), 0, s,
                       p->offsets.as<uint64_t>(), p->block_sums.as<uint64_t>(), &g->n);
    hipLaunchKernelGGL(k_rep_sites, dim3(1), dim3(64), 0, s, g, p->offsets.as<uint64_t>());
    const unsigned gcap = (unsigned)std::min<uint64_t>((cap + 255) / 256, 2048);
    hipLaunchKernelGGL(k_rep_order, dim3(gcap), dim3(256), 0, s, p->stage.as<abx_banked>(),
                       p->stage_parent.as<uint32_t>(), p->stage_daughter.as<uint32_t>(),
                       p->offsets.as<uint64_t>(), g, p->ordered.as<abx_banked>());
    hipLaunchKernelGGL(k_rep_end, dim3(1), dim3(64), 0, s, g, W, conv, hist_dev, khist_dev);
    hipLaunchKernelGGL(k_rebank_rep, dim3(gcap), dim3(256), 0, s, p->ordered.as<abx_banked>(),
                       g, T.st.stride, out);
    if (conv && tb > 0)
      hipLaunchKernelGGL(k_record_rep, dim3((unsigned)blocks_for(tb, 256)), dim3(256), 0, s,
                         p->tally_gen.as<double>(), p->tally_avg.as<double>(),
                         p->tally_var.as<double>(), tb, g, 1.0 / (double)R);
    HIPC(hipStreamEndCapture(s, &gr[par][conv][1]));
    for (int k = 0; k < 2; k++)
      HIPC(hipGraphInstantiate(&ex[par][conv][k], gr[par][conv][k], nullptr, nullptr, 0));
  };
  // (destroys each graph once: an error after the first call lands in the
  // catch below, which calls it again)
  auto cleanup = [&]() {
    for (auto& a : ex)
      for (auto& b : a)
        for (auto& c : b)
          if (c) {
            (void)hipGraphExecDestroy(c);
            c = nullptr;
          }
    for (auto& a : gr)
      for (auto& b : a)
        for (auto& c : b)
          if (c) {
            (void)hipGraphDestroy(c);
            c = nullptr;
          }
  };
  try {
    while ((int)p->gen_events.size() < 2 * ngenerations + 2) {
      hipEvent_t e;
      HIPC(hipEventCreate(&e));
      p->gen_events.push_back(e);
    }
    bool converged = nignored == 0;
    int par = 0;
    int queued = ngenerations;  // generations launched (fewer after a stop)
    for (int gi = 0; gi < ngenerations; gi++) {
      if (gi >= 2) {
        HIPC(hipEventSynchronize(p->gen_events[2 * (gi - 1)]));
        if (stop_now(p, p->run_gens + gi - 1, 1)) {
          queued = gi;
          break;
        }
      }
      const int conv = converged ? 1 : 0;
      if (!ex[par][conv][0]) capture(par, conv);
      HIPC(hipEventRecord(p->gen_events[2 * gi], s));
      HIPC(hipGraphLaunch(ex[par][conv][0], s));
      HIPC(hipEventRecord(p->gen_events[2 * gi + 1], s));
      HIPC(hipGraphLaunch(ex[par][conv][1], s));
      par ^= 1;
      if (gi + 1 == nignored) converged = true;
    }
    HIPC(hipEventRecord(p->gen_events[2 * queued], s));
    HIPC(hipMemcpyAsync(&h, g, sizeof(GenRep), hipMemcpyDeviceToHost, s));
    std::vector<double> hist((size_t)(6 + R) * ngenerations);
    HIPC(hipMemcpyAsync(hist.data(), hist_dev, hist.size() * sizeof(double),
                        hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    float total = 0.f;
    double kms = 0.0;
    for (int gi = 0; gi < queued; gi++) {
      float x = 0.f;
      HIPC(hipEventElapsedTime(&x, p->gen_events[2 * gi], p->gen_events[2 * gi + 1]));
      kms += x;
    }
    HIPC(hipEventElapsedTime(&total, p->gen_events[0], p->gen_events[2 * queued]));
    cleanup();
    if (h.err) {
      if (h.err == -8) fail(ABX_E_INVALID, "No fission neutrons were produced.");
      fail(h.err, "transport failed at history " + std::to_string(h.err_hist));
    }
    p->gen = h.gen;
    p->tally_g = h.gen;
    for (int q = 0; q < 6; q++) {
      p->s_avg[q] = h.s_avg[q];
      p->s_var[q] = h.s_var[q];
    }
    const int ran = std::min(queued, (int)h.hist_idx);
    p->run_gens += ran;
    for (int q = 0; q < 6; q++)
      p->gen_vals[q].insert(p->gen_vals[q].end(), hist.begin() + (size_t)q * ngenerations,
                            hist.begin() + (size_t)q * ngenerations + ran);
    if (kcol_hist)
      std::copy(hist.begin() + (size_t)6 * ngenerations,
                hist.begin() + (size_t)(6 + R) * ngenerations, kcol_hist);
    p->gen_scores = abx_scores{};
    p->run_started = false;  // the next abx_run starts afresh
    p->bank_n = 0;
    p->last_sites = 0;
    // the latest bank is in bank_b after an odd number of generations
    if (par) std::swap(p->bank_a, p->bank_b);
    p->rep_chains = R;
    *stats = abx_run_stats{};
    fill_statistics(p, stats);
    abx_diag& d = stats->diag;
    d.histories = h.diag[0];
    d.collisions = h.diag[1];
    d.virtual_collisions = h.diag[2];
    d.flights = h.diag[3];
    d.crossings = h.diag[4];
    d.tally_updates = h.diag[5];
    d.lost = h.diag[6];
    d.secondaries = h.diag[7];
    stats->transport_seconds = total * 1e-3;
    stats->kernel_seconds = kms * 1e-3;
    stats->kernel_launches = queued;
    stats->generations_run = ran;
    stats->kernel_items = (int64_t)(d.flights + d.collisions);
    stats->histories = (int64_t)d.histories;
  } catch (...) {
    cleanup();
    throw;
  }
}

// FixedSource::run / ModifiedFixedSource::run / PowerIterator::run
// (src/fixed_source.cpp:81-173, src/modified_fixed_source.cpp:59-143,
// src/power_iterator.cpp:305-429) with every bank resident in HBM.
static void run_generations(abx_problem* p, int ngenerations, int nignored, double* kcol_hist,
                            abx_run_stats* stats) {
  const HostTables& T = p->T;
  const int64_t N = T.st.nparticles;
  const double W = (double)N;
  *stats = abx_run_stats{};
  bool converged = T.st.mode != ABX_K_EIGENVALUE || nignored == 0;
  p->timing = true;
  p->kernel_ms = 0.0;
  p->kernel_launches = 0;
  p->kernel_items = 0;
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  float ms_total = 0.f;
  int64_t histories = 0;
  // bank_a holds the current k-eigenvalue source between calls: it is only
  // (re)allocated when sources are sampled into it.
  if (T.st.mode != ABX_K_EIGENVALUE) {
    p->bank_a.reserve((size_t)N * sizeof(abx_particle));
  } else if (!p->run_started) {
    p->bank_a.reserve((size_t)N * sizeof(abx_particle));
    sample_sources_device(p, 0, N, p->bank_a.as<abx_particle>());
    p->run_hist = (uint64_t)N;
    p->bank_n = (size_t)N;
  }
  p->run_started = true;
  if (run_generations_graph(p, ngenerations, nignored, kcol_hist, stats)) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->timing = false;
    return;
  }
  int ran = 0;
  for (int g = 1; g <= ngenerations; g++) {
    abx_diag dg{};
    abx_scores sc{};
    if (T.st.mode != ABX_K_EIGENVALUE) {
      // A modified-fixed-source chain swaps bank_a with bank_b, which was
      // sized for the chain's last re-banked generation: the next batch's
      // N sources need bank_a grown back to N records.
      p->bank_a.reserve((size_t)N * sizeof(abx_particle));
      require_records(p->bank_a, N, sizeof(abx_particle), "source bank");
      sample_sources_device(p, p->run_hist, N, p->bank_a.as<abx_particle>());
      p->run_hist += (uint64_t)N;
      int64_t n = N;
      HIPC(hipEventRecord(e0, p->stream));
      TransportOut r = transport_device(p, p->bank_a.as<abx_particle>(), n, p->run_kcol, 1, nullptr, 0, nullptr);
      histories += n;
      add_diag(dg, r.diag);
      sc = r.scores;
      while (T.st.mode == ABX_MODIFIED_FIXED_SOURCE && r.n_sites > 0) {
        n = r.n_sites;
        if (p->cancel) cancel_device(p, p->ordered.as<abx_banked>(), n);
        p->bank_b.reserve((size_t)n * sizeof(abx_particle));
        require_records(p->bank_b, n, sizeof(abx_particle), "re-banked generation");
        hipLaunchKernelGGL(k_rebank, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, p->stream,
                           p->ordered.as<abx_banked>(), n, 1.0, p->run_hist, T.st.seed, T.st.stride, 1, 0,
                           p->bank_b.as<abx_particle>());
        HIPC(hipGetLastError());
        p->run_hist += (uint64_t)n;
        std::swap(p->bank_a, p->bank_b);
        r = transport_device(p, p->bank_a.as<abx_particle>(), n, p->run_kcol, 1, nullptr, 0, nullptr);
        histories += n;
        add_diag(dg, r.diag);
        add_scores(sc, r.scores);
      }
      HIPC(hipEventRecord(e1, p->stream));
      HIPC(hipEventSynchronize(e1));
      float ms = 0;
      HIPC(hipEventElapsedTime(&ms, e0, e1));
      ms_total += ms;
      if (kcol_hist) kcol_hist[g - 1] = sc.k_col / W;
      push_gen_values(p, sc, W);
      record_generation(p, sc, W);
    } else {
      HIPC(hipEventRecord(e0, p->stream));
      TransportOut r = transport_device(p, p->bank_a.as<abx_particle>(), (int64_t)p->bank_n, p->run_kcol,
                                        converged ? 1 : 0, nullptr, 0, nullptr);
      HIPC(hipEventRecord(e1, p->stream));
      HIPC(hipEventSynchronize(e1));
      float ms = 0;
      HIPC(hipEventElapsedTime(&ms, e0, e1));
      ms_total += ms;
      histories += (int64_t)p->bank_n;
      add_diag(dg, r.diag);
      sc = r.scores;
      if (r.n_sites == 0) fail(ABX_E_INVALID, "No fission neutrons were produced.");
      const double kc = sc.k_col / W;
      if (kcol_hist) kcol_hist[g - 1] = kc;
      push_gen_values(p, sc, W);
      // perform_regional_cancellation, then normalize_weights: W = W_pos -
      // W_neg, w *= nparticles / W (src/power_iterator.cpp:354-359)
      if (p->cancel) cancel_device(p, p->ordered.as<abx_banked>(), r.n_sites);
      if (p->exact) r.n_sites = exact_cancel_device(p, r.n_sites);
      p->wsum.reserve(2 * sizeof(double));
      HIPC(hipMemsetAsync(p->wsum.p, 0, 2 * sizeof(double), p->stream));
      hipLaunchKernelGGL(k_weight_sums, dim3((unsigned)std::min<int64_t>(1024, blocks_for(r.n_sites, 256))), dim3(256), 0,
                         p->stream, p->ordered.as<abx_banked>(), r.n_sites, p->wsum.as<double>());
      HIPC(hipGetLastError());
      double ws[2];
      HIPC(hipMemcpyAsync(ws, p->wsum.p, 2 * sizeof(double), hipMemcpyDeviceToHost, p->stream));
      HIPC(hipStreamSynchronize(p->stream));
      const double w_per_part = (double)T.st.nparticles / (ws[0] - ws[1]);
      if (converged) record_generation(p, sc, W);
      p->bank_b.reserve((size_t)r.n_sites * sizeof(abx_particle));
      hipLaunchKernelGGL(k_rebank, dim3((unsigned)blocks_for(r.n_sites, 256)), dim3(256), 0,
                         p->stream, p->ordered.as<abx_banked>(), r.n_sites, w_per_part,
                         p->run_hist, T.st.seed, T.st.stride, 0, 1, p->bank_b.as<abx_particle>());
      HIPC(hipGetLastError());
      std::swap(p->bank_a, p->bank_b);
      p->bank_n = (size_t)r.n_sites;
      p->run_hist += (uint64_t)r.n_sites;
      p->run_kcol = kc;
      if (g == nignored) converged = true;
    }
    clear_generation(p);
    add_diag(stats->diag, dg);
    p->run_gens++;
    ran = g;
    if (g < ngenerations && stop_now(p, p->run_gens, 0)) break;
  }
  HIPC(hipStreamSynchronize(p->stream));
  p->timing = false;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  fill_statistics(p, stats);
  stats->generations_run = ran;
  stats->transport_seconds = ms_total * 1e-3;
  stats->kernel_seconds = p->kernel_ms * 1e-3;
  stats->kernel_launches = p->kernel_launches;
  stats->kernel_items = p->engine == ABX_ENGINE_HISTORY
                            ? (int64_t)(stats->diag.flights + stats->diag.collisions)
                            : p->kernel_items;
  stats->histories = histories;
}

int abx_request_stop(abx_problem* p) {
  if (!p) return ABX_E_INVALID;
  p->stop_requested.store(1);
  return ABX_OK;
}

int abx_set_max_time(abx_problem* p, double seconds) {
  ABX_TRY({
    if (!p || !(seconds == seconds)) fail(ABX_E_INVALID, "max time");
    p->max_time = seconds;
  });
}

int abx_run(abx_problem* p, int ngenerations, int nignored, double* kcol_hist,
            abx_run_stats* stats) {
  ABX_TRY({
    if (!p || !stats || ngenerations < 1 || nignored < 0) fail(ABX_E_INVALID, "run arguments");
    run_reset(p);
    run_generations(p, ngenerations, nignored, kcol_hist, stats);
  });
}

int abx_run_continue(abx_problem* p, int ngenerations, int nignored, double* kcol_hist,
                     abx_run_stats* stats) {
  ABX_TRY({
    if (!p || !stats || ngenerations < 1 || nignored < 0) fail(ABX_E_INVALID, "run arguments");
    run_generations(p, ngenerations, nignored, kcol_hist, stats);
  });
}

int abx_run_replicas(abx_problem* p, int replicas, const uint64_t* seeds, int from_bank,
                     int ngenerations, int nignored, double* kcol_hist, abx_run_stats* stats) {
  ABX_TRY({
    if (!p || !stats || ngenerations < 1 || nignored < 0) fail(ABX_E_INVALID, "run arguments");
    if (replicas < 1 || replicas > kMaxReplicas)
      fail(ABX_E_INVALID, "replicas must be 1.." + std::to_string(kMaxReplicas));
    run_replicas(p, replicas, seeds, from_bank != 0, ngenerations, nignored, kcol_hist, stats,
                 false);
  });
}

int abx_run_replicas_continue(abx_problem* p, int ngenerations, int nignored,
                              double* kcol_hist, abx_run_stats* stats) {
  ABX_TRY({
    if (!p || !stats || ngenerations < 1 || nignored < 0) fail(ABX_E_INVALID, "run arguments");
    run_replicas(p, p->rep_chains, nullptr, false, ngenerations, nignored, kcol_hist, stats,
                 true);
  });
}

int abx_gen_values(abx_problem* p, int quantity, double* out, size_t cap, size_t* n) {
  ABX_TRY({
    if (!p || quantity < 0 || quantity > 5) fail(ABX_E_INVALID, "quantity");
    const std::vector<double>& v = p->gen_vals[quantity];
    if (n) *n = v.size();
    if (out && !v.empty()) {  // out == NULL: the count only
      if (v.size() > cap) fail(ABX_E_CAPACITY, "per-generation buffer too small");
      std::copy(v.begin(), v.end(), out);
    }
  });
}

int abx_cancel_bank(abx_problem* p, abx_banked* bank, size_t n) {
  ABX_TRY({
    if (!p || (!bank && n)) fail(ABX_E_INVALID, "null argument");
    if (!p->cancel) fail(ABX_E_INVALID, "no cancellation mesh in the description");
    if (n > 0xffffffffull) fail(ABX_E_INVALID, "bank larger than 2^32");
    p->ordered.reserve(std::max<size_t>(1, n) * sizeof(abx_banked));
    if (n) HIPC(hipMemcpyAsync(p->ordered.p, bank, n * sizeof(abx_banked), hipMemcpyHostToDevice, p->stream));
    cancel_device(p, p->ordered.as<abx_banked>(), (int64_t)n);
    if (n) HIPC(hipMemcpyAsync(bank, p->ordered.p, n * sizeof(abx_banked), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_parent_fetch(abx_problem* p, abx_parent_info* out, size_t cap, size_t* n) {
  ABX_TRY({
    if (!p || !n || (!out && cap)) fail(ABX_E_INVALID, "null argument");
    if (!p->exact) fail(ABX_E_INVALID, "no exact cancelator in the description");
    const size_t m = (size_t)p->last_sites;
    *n = m;
    if (m > cap) fail(ABX_E_CAPACITY, "parent buffer holds " + std::to_string(cap) + " records, " +
                                          std::to_string(m) + " needed");
    if (m) {
      require_records(p->ordered_pinfo, (int64_t)m, sizeof(abx_parent_info), "parent records");
      HIPC(hipMemcpyAsync(out, p->ordered_pinfo.p, m * sizeof(abx_parent_info), hipMemcpyDeviceToHost,
                          p->stream));
    }
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_exact_cancel_bank(abx_problem* p, const abx_banked* bank, const abx_parent_info* parents,
                          size_t n, abx_banked* out, size_t cap, size_t* n_out) {
  ABX_TRY({
    if (!p || !n_out || (n && (!bank || !parents)) || (!out && cap))
      fail(ABX_E_INVALID, "null argument");
    if (!p->exact) fail(ABX_E_INVALID, "no exact cancelator in the description");
    if (n > 0xffffffffull) fail(ABX_E_INVALID, "bank larger than 2^32");
    p->ordered.reserve(std::max<size_t>(1, n) * sizeof(abx_banked));
    p->ordered_pinfo.reserve(std::max<size_t>(1, n) * sizeof(abx_parent_info));
    if (n) {
      HIPC(hipMemcpyAsync(p->ordered.p, bank, n * sizeof(abx_banked), hipMemcpyHostToDevice, p->stream));
      HIPC(hipMemcpyAsync(p->ordered_pinfo.p, parents, n * sizeof(abx_parent_info),
                          hipMemcpyHostToDevice, p->stream));
    }
    p->last_sites = 0;  // `ordered` no longer holds the last transport's sites
    const int64_t m = exact_cancel_device(p, (int64_t)n);
    *n_out = (size_t)m;
    if ((size_t)m > cap)
      fail(ABX_E_CAPACITY, "out holds " + std::to_string(cap) + " records, " + std::to_string(m) +
                               " needed");
    if (m) HIPC(hipMemcpyAsync(out, p->ordered.p, (size_t)m * sizeof(abx_banked), hipMemcpyDeviceToHost,
                               p->stream));
    HIPC(hipStreamSynchronize(p->stream));
  });
}

int abx_rng_reset(abx_problem* p) {
  ABX_TRY({
    if (!p) fail(ABX_E_INVALID, "null argument");
    p->grng = pcg_seed(p->T.st.seed);
  });
}

int abx_rng_get_state(abx_problem* p, uint64_t* state) {
  ABX_TRY({
    if (!p || !state) fail(ABX_E_INVALID, "null argument");
    *state = p->grng;
  });
}

int abx_rng_set_state(abx_problem* p, uint64_t state) {
  ABX_TRY({
    if (!p) fail(ABX_E_INVALID, "null argument");
    p->grng = state;
  });
}

int abx_bank_size(abx_problem* p, size_t* n) {
  ABX_TRY({
    if (!p || !n) fail(ABX_E_INVALID, "null argument");
    *n = p->bank_n;
  });
}

int abx_bank_load(abx_problem* p, const double* in9, size_t n) {
  ABX_TRY({
    if (!p || (!in9 && n)) fail(ABX_E_INVALID, "null argument");
    if (p->T.st.mode != ABX_K_EIGENVALUE) fail(ABX_E_INVALID, "a source bank seeds k-eigenvalue runs");
    if (n == 0) fail(ABX_E_INVALID, "empty source bank");
    std::vector<abx_particle> b(n);
    double tot = 0.0;
    for (size_t i = 0; i < n; i++) {
      const double* s9 = in9 + 9 * i;
      abx_particle& a = b[i];
      const double m = std::sqrt(s9[3] * s9[3] + s9[4] * s9[4] + s9[5] * s9[5]);
      for (int k = 0; k < 3; k++) {
        a.r[k] = s9[k];
        a.u[k] = s9[3 + k] / m;  // Direction's normalising constructor
      }
      a.E = s9[6];
      a.wgt = s9[7];
      a.wgt2 = s9[8];
      a.history_id = i;
      a.family_id = 0;
      a.rng_state = pcg_history_state(p->T.st.seed, p->T.st.stride, i);
      tot += s9[7];
    }
    run_reset(p);
    p->bank_a.reserve(n * sizeof(abx_particle));
    HIPC(hipMemcpyAsync(p->bank_a.p, b.data(), n * sizeof(abx_particle), hipMemcpyHostToDevice, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    p->bank_n = n;
    p->run_hist = n;
    p->run_started = true;
    p->T.st.nparticles = std::max<int64_t>(1, std::llround(tot));
  });
}

int abx_bank_fetch(abx_problem* p, double* out9) {
  ABX_TRY({
    if (!p || !out9) fail(ABX_E_INVALID, "null argument");
    std::vector<abx_particle> b(p->bank_n);
    if (p->bank_n)
      HIPC(hipMemcpyAsync(b.data(), p->bank_a.p, p->bank_n * sizeof(abx_particle), hipMemcpyDeviceToHost,
                          p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    for (size_t i = 0; i < b.size(); i++) {
      double* o = out9 + 9 * i;
      for (int k = 0; k < 3; k++) {
        o[k] = b[i].r[k];
        o[3 + k] = b[i].u[k];
      }
      o[6] = b[i].E;
      o[7] = b[i].wgt;
      o[8] = b[i].wgt2;
    }
  });
}

}  // extern "C"

// ============================================================== batches
// Many fixed-source / modified-fixed-source problems advanced together: one
// sampling launch per problem, then ONE persistent transport launch per
// kernel variant (tracking x secondaries x precision) over the histories of
// every attached problem, then per-problem bookkeeping.  A problem's results
// are those of abx_run on it alone (same histories, same RNG streams, same
// ordered banks); only the summation order of tallies and scores differs.

namespace {

struct BatchEntry {
  abx_problem* p;
  int batches_left;
  bool mid_chain;     // a modified-fixed-source generation chain is running
  int64_t n;          // histories of the current call
  abx_scores acc{};   // scores of the current batch (chain or slices summed)
  int64_t done = 0;   // histories of the current fixed-source batch already run
  size_t foot = 0;    // device bytes counted against the batch's budget
};

// A fixed-source batch may be split over several steps (its histories are
// independent and its tallies accumulate until the batch is recorded); a
// modified-fixed-source generation chain runs one whole generation per step.
bool sliceable(const abx_problem* p) { return p->T.st.mode == ABX_FIXED_SOURCE; }

// Histories each entry runs this step.  With a step budget, entries whose
// work cannot be split take what they need and the sliceable ones share the
// rest by water-filling (smallest remainder first), each getting at least
// `floor` histories so that every problem advances.
std::vector<int64_t> step_allocation(const std::vector<BatchEntry>& items, int64_t budget) {
  const int ni = (int)items.size();
  const int64_t floor = 8192;
  std::vector<int64_t> give(ni, 0), idx;
  int64_t left = budget;
  for (int j = 0; j < ni; j++) {
    const BatchEntry& it = items[j];
    const int64_t N = it.p->T.st.nparticles;
    if (it.mid_chain) {
      give[j] = it.n;
    } else if (!sliceable(it.p)) {
      give[j] = N;
    } else {
      give[j] = N - it.done;
      idx.push_back(j);
      continue;
    }
    left -= give[j];
  }
  if (budget <= 0) return give;
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t c) { return give[a] < give[c]; });
  int64_t k = (int64_t)idx.size();
  for (int64_t j : idx) {
    const int64_t share = std::max<int64_t>(floor, left / std::max<int64_t>(1, k));
    give[j] = std::min(give[j], share);
    left -= give[j];
    k--;
  }
  return give;
}

// Per-entry device scalars in the batch-wide block.
struct BatchScalars {
  double sc[8];
  unsigned long long cn[8];
  unsigned long long err_hist;
  unsigned long long n_stage;
  unsigned long long n_hole;
  int32_t err;
  int32_t pad_;
};

// Whether the batch kernel would run this problem with its tables staged in
// LDS: they fit next to 256 cursors and cost no resident block.
bool batch_stage_fits(const abx_problem* p) {
  const HostTables& T = p->T;
  const int trk = T.st.tracking;
  const bool sec = needs_secondaries(T), split = T.st.precision == ABX_SPLIT;
  const size_t cur = cursor_lds_bytes(256, p->P.trk_D, p->P.trk_L);
  bool ok = false;
  (void)dispatch_variant(trk, false, sec || trk == 2, split, T.extended,
                         [&](auto TT, auto, auto SE, auto SP, auto EX) {
    constexpr int TK = decltype(TT)::value;
    constexpr bool SEC = decltype(SE)::value, SPL = decltype(SP)::value;
    constexpr bool EXT = decltype(EX)::value;
    ok = stage_pays(k_batch_chunks<TK, SEC, SPL, true, EXT>,
                    k_batch_chunks<TK, SEC, SPL, false, EXT>, 256, cur, (size_t)p->P.blob_bytes);
    return hipSuccess;
  });
  return ok;
}

// Launch group of a problem in the batch kernel: tracking mode, secondaries,
// precision, extended geometry, and whether its tables are worth staging in
// LDS (problems whose tables are launched apart from those whose tables are
// not, so that the former run the staged variant).
int variant_key(abx_problem* p) {
  const HostTables& T = p->T;
  const int sec = needs_secondaries(T) ? 1 : 0;
  const int split = T.st.precision == ABX_SPLIT ? 1 : 0;
  const int ext = T.extended ? 1 : 0;
  if (p->batch_stage < 0) p->batch_stage = batch_stage_fits(p) ? 1 : 0;
  return (((T.st.tracking * 2 + sec) * 2 + split) * 2 + ext) * 2 + p->batch_stage;
}

// Chunks of `chunk` histories over the problems of one launch (item indices
// local to the launch's arrays).
std::vector<BatchChunk> make_chunks(const std::vector<int64_t>& n, int64_t chunk) {
  std::vector<BatchChunk> c;
  for (size_t j = 0; j < n.size(); j++)
    for (int64_t a = 0; a < n[j]; a += chunk)
      c.push_back(BatchChunk{(int32_t)j, 0, a, std::min(n[j], a + chunk)});
  return c;
}

// Chunk length: about `per_cu` chunks per CU over the launch, 256 - 8192
// histories (at least one block's lanes, so that lanes <= histories + 255).
int64_t chunk_length(int64_t total, int per_cu = 8) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  return std::max<int64_t>(256, std::min<int64_t>(8192, total / std::max(1, per_cu * cus)));
}

// The chunk counter `next` was zeroed by the caller.
template <int TRK, bool SEC, bool SPLIT, bool STAGE, bool EXT>
hipError_t launch_chunks(const DevProblem* probs, const KCtx* ctxs, const BatchChunk* chunks,
                         int n_chunks, int D, int L, size_t blob_max, unsigned* next, int bs,
                         hipStream_t s) {
  const size_t lds = cursor_lds_bytes(bs, D, L) + (STAGE ? blob_max : 0);
  const int64_t blocks = std::min<int64_t>(
      n_chunks, resident_blocks(k_batch_chunks<TRK, SEC, SPLIT, STAGE, EXT>, bs, lds));
  hipLaunchKernelGGL((k_batch_chunks<TRK, SEC, SPLIT, STAGE, EXT>), dim3((unsigned)blocks),
                     dim3(bs), lds, s, probs, ctxs, chunks, n_chunks, D, L, next);
  return hipGetLastError();
}

// One launch over the chunks of problems of one kernel variant, blocks of
// bs lanes; the tables are staged in LDS when the largest blob fits next to
// the cursors.
hipError_t launch_chunks_variant(int key, const DevProblem* probs, const KCtx* ctxs,
                                 const BatchChunk* chunks, int n_chunks, int D, int L,
                                 size_t blob_max, unsigned* next, int bs, hipStream_t s) {
  const int trk = key / 16;
  const bool sec = (key / 8) & 1, split = (key / 4) & 1, ext = (key / 2) & 1, fits = key & 1;
  const size_t cur = cursor_lds_bytes(bs, D, L);
  return dispatch_variant(trk, false, sec || trk == 2, split, ext,
                          [&](auto T, auto, auto SE, auto SP, auto EX) {
    constexpr int TK = decltype(T)::value;
    constexpr bool SEC = decltype(SE)::value, SPL = decltype(SP)::value;
    constexpr bool EXT = decltype(EX)::value;
    if (fits && stage_pays(k_batch_chunks<TK, SEC, SPL, true, EXT>,
                           k_batch_chunks<TK, SEC, SPL, false, EXT>, bs, cur, blob_max))
      return launch_chunks<TK, SEC, SPL, true, EXT>(probs, ctxs, chunks, n_chunks, D, L, blob_max,
                                                    next, bs, s);
    return launch_chunks<TK, SEC, SPL, false, EXT>(probs, ctxs, chunks, n_chunks, D, L, blob_max,
                                                   next, bs, s);
  });
}

// Bytes a problem holds on the device while it runs (tables, tallies,
// banks, fission staging) - what admission checks against a budget.
size_t problem_footprint(const abx_problem* p) {
  const HostTables& T = p->T;
  const size_t N = (size_t)T.st.nparticles;
  size_t b = 0;
  for (const auto& a : p->arrays) b += a.bytes;
  b += 3 * (size_t)std::max<int64_t>(1, T.tally_bins) * sizeof(double);
  b += 2 * N * sizeof(abx_particle);
  if (T.st.mode != ABX_FIXED_SOURCE)
    b += (4 * N + 4096) * (sizeof(abx_banked) + 2 * sizeof(uint32_t)) +
         N * (sizeof(uint32_t) + sizeof(uint64_t) + sizeof(abx_banked));
  return b;
}

}  // namespace

struct abx_batch {
  std::vector<BatchEntry> items;
  hipStream_t stream = nullptr;
  DBuf probs, ctxs, offsets, scal, next, sec, chunks;
  std::vector<BatchScalars> h_scal;
  // host images of a step's launch arrays (read by the uploads)
  std::vector<DevProblem> hp;
  std::vector<KCtx> hk;
  std::vector<BatchChunk> all_chunks;
  std::vector<SampleJob> sjobs;
  std::vector<DevSources> ssrc;
  std::vector<RecordJob> rjobs;
  DBuf sjob_d, ssrc_d, rjob_d;
  size_t budget = 0;  // device bytes admitted problems may hold (0 = no limit)
  size_t used = 0;
  int64_t step_hist = 0;  // histories per step over all problems (0 = no limit)
  int64_t chunk = 0;      // histories per chunk (0 = about 8 per CU, -f = about f per CU)
  int64_t multi_hist = int64_t(1) << 23;  // cap on side-by-side batches of one problem
  double kernel_ms = 0.0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // Launch groups (kernel variants) run on these streams side by side, so
  // that one group's tail (its last, longest chunks) overlaps the next
  // group's work.
  static constexpr int NS = 4;
  hipStream_t aux[NS] = {};
  hipEvent_t ready = nullptr, done[NS] = {};
  int block = 256;  // lanes per block of the batch kernel (64, 128 or 256)
  // A step whose launches have the previous step's shape (kernel variants,
  // chunk counts, block, device arrays) replays the transport captured as a
  // graph instead of launching it again.
  bool graphs = true;
  std::vector<int64_t> last_sig;  // launch signature of the previous step
  hipGraphExec_t gexec = nullptr;       // its captured transport, or none
  int64_t captures = 0, replays = 0;
  ~abx_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    DBuf* all[] = {&probs, &ctxs, &offsets, &scal, &next, &sec, &chunks, &sjob_d, &ssrc_d, &rjob_d};
    for (DBuf* b : all) b->release();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (ready) (void)hipEventDestroy(ready);
    for (int k = 0; k < NS; k++) {
      if (done[k]) (void)hipEventDestroy(done[k]);
      if (aux[k]) (void)hipStreamDestroy(aux[k]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Concurrent batches of one fixed-source problem in a step: up to
// kMultiMax, and at most multi_hist histories together (a problem with
// large batches runs one at a time).
constexpr int kMultiMax = 8;

int concurrent_batches(const abx_problem* p, int batches_left, int64_t step_hist,
                       int64_t multi_hist) {
  const int64_t N = p->T.st.nparticles;
  if (!sliceable(p) || step_hist > 0 || N <= 0) return 1;
  const int64_t k = std::min<int64_t>({(int64_t)batches_left, (int64_t)kMultiMax,
                                       std::max<int64_t>(1, multi_hist / N)});
  return (int)std::max<int64_t>(1, k);
}

// One launch item: histories [0, n) of bank `in` scoring into `gen`.
struct LaunchItem {
  int entry;
  int k;  // concurrent batch index within the entry
  int64_t n;
  const abx_particle* in;
  double* gen;
};

// Device bytes a problem holds while attached to b for `nbatches` batches:
// the problem alone, plus the banks and generation arrays of the batches it
// runs side by side.
size_t batch_need(const abx_batch* b, const abx_problem* p, int nbatches) {
  const int kb = concurrent_batches(p, nbatches, b->step_hist, b->multi_hist);
  size_t need = problem_footprint(p);
  if (kb > 1)
    need += (size_t)(kb - 1) * ((size_t)p->T.st.nparticles * sizeof(abx_particle) +
                                (size_t)std::max<int64_t>(1, p->T.tally_bins) * sizeof(double));
  return need;
}

void batch_step(abx_batch* b, std::vector<abx_problem*>& finished) {
  const int ne = (int)b->items.size();
  if (ne == 0) return;
  // 1. launch items and their sources: the next batches (several at once
  // for small fixed-source problems, or this step's slice of a large one),
  // or the chain's re-banked generation
  const std::vector<int64_t> give = step_allocation(b->items, b->step_hist);
  std::vector<LaunchItem> li;
  std::vector<int> first_item(ne);
  for (int j = 0; j < ne; j++) {
    BatchEntry& it = b->items[j];
    abx_problem* p = it.p;
    first_item[j] = (int)li.size();
    if (it.mid_chain) {
      li.push_back(LaunchItem{j, 0, it.n, p->bank_a.as<abx_particle>(), p->tally_gen.as<double>()});
      continue;
    }
    const int kb = concurrent_batches(p, it.batches_left, b->step_hist, b->multi_hist);
    const int64_t n = kb > 1 ? p->T.st.nparticles : give[j];
    const int64_t nb = std::max<int64_t>(1, p->T.tally_bins);
    p->bank_a.reserve((size_t)std::max<int64_t>(1, kb * n) * sizeof(abx_particle));
    if (kb > 1 && p->tally_more.bytes < (size_t)(kb - 1) * nb * sizeof(double)) {
      // generation arrays are zero between steps; a new one starts zeroed
      p->tally_more.reserve((size_t)(kb - 1) * nb * sizeof(double));
      hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks_for((kb - 1) * nb, 256)), dim3(256), 0,
                         b->stream, p->tally_more.as<double>(), (int64_t)(kb - 1) * nb, 0.0);
      HIPC(hipGetLastError());
    }
    for (int k = 0; k < kb; k++)
      li.push_back(LaunchItem{j, k, n, p->bank_a.as<abx_particle>() + (size_t)k * n,
                              k == 0 ? p->tally_gen.as<double>()
                                     : p->tally_more.as<double>() + (size_t)(k - 1) * nb});
    if (it.done == 0) it.acc = abx_scores{};
    it.n = n;
  }
  const int ni = (int)li.size();
  b->scal.reserve((size_t)ni * sizeof(BatchScalars));
  HIPC(hipMemsetAsync(b->scal.p, 0, (size_t)ni * sizeof(BatchScalars), b->stream));
  BatchScalars* ds = b->scal.as<BatchScalars>();
  // 2. one launch per kernel variant
  std::vector<int> order(ni);
  for (int i = 0; i < ni; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    return variant_key(b->items[li[a].entry].p) < variant_key(b->items[li[c].entry].p);
  });
  int max_sec = 0;
  for (auto& it : b->items) max_sec = std::max(max_sec, needs_secondaries(it.p->T) ? 16 : 0);
  int dev = 0, cus = 0;
  HIPC(hipGetDevice(&dev));
  HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const size_t lanes = (size_t)cus * 8 * 256;  // upper bound on resident lanes
  // one secondary stack per stream slot: groups on different streams run
  // at the same time and index their stacks by lane
  const int NS = abx_batch::NS;
  if (max_sec) b->sec.reserve((size_t)NS * lanes * max_sec * sizeof(SecRec));
  b->hp.resize(ni);
  b->hk.resize(ni);
  b->probs.reserve((size_t)ni * sizeof(DevProblem));
  b->ctxs.reserve((size_t)ni * sizeof(KCtx));
  if (!b->e0) {
    HIPC(hipEventCreate(&b->e0));
    HIPC(hipEventCreate(&b->e1));
    HIPC(hipEventCreateWithFlags(&b->ready, hipEventDisableTiming));
    for (int k = 0; k < NS; k++) {
      HIPC(hipStreamCreateWithFlags(&b->aux[k], hipStreamNonBlocking));
      HIPC(hipEventCreateWithFlags(&b->done[k], hipEventDisableTiming));
    }
  }
  // chunks of every variant group, uploaded once
  struct Group {
    int key, g0, cnt, D, L;
    size_t blob_max;
    int64_t c0, nc;
  };
  std::vector<Group> groups;
  std::vector<int> pos(ni);  // item -> its position in the launch arrays
  b->all_chunks.clear();
  int64_t total_all = 0;
  for (const auto& x : li) total_all += x.n;
  const int64_t clen = b->chunk > 0   ? std::max<int64_t>(64, b->chunk)
                       : b->chunk < 0 ? chunk_length(total_all, (int)std::min<int64_t>(1024, -b->chunk))
                                      : chunk_length(total_all);
  int g0 = 0;
  while (g0 < ni) {
    const int key = variant_key(b->items[li[order[g0]].entry].p);
    int g1 = g0;
    Group gr{key, g0, 0, 1, 1, 16, (int64_t)b->all_chunks.size(), 0};
    std::vector<int64_t> n;
    while (g1 < ni && variant_key(b->items[li[order[g1]].entry].p) == key) {
      const int i = order[g1];
      const LaunchItem& x = li[i];
      abx_problem* p = b->items[x.entry].p;
      const CallScalars cs{ds[i].sc, ds[i].cn, &ds[i].err, &ds[i].err_hist, &ds[i].n_stage,
                           &ds[i].n_hole};
      const int sec_cap = needs_secondaries(p->T) ? 16 : 0;
      SecRec* sec = max_sec ? b->sec.as<SecRec>() + (groups.size() % NS) * lanes * max_sec
                            : nullptr;
      b->hk[g1] = make_kctx(p, x.in, x.n, p->run_kcol, 1, cs, sec, sec_cap);
      b->hp[g1] = p->P;
      pos[i] = g1;
      b->hp[g1].tally_data = x.gen;
      n.push_back(x.n);
      gr.D = std::max(gr.D, p->P.trk_D);
      gr.L = std::max(gr.L, p->P.trk_L);
      gr.blob_max = std::max(gr.blob_max, (size_t)p->P.blob_bytes);
      g1++;
    }
    gr.cnt = g1 - g0;
    const std::vector<BatchChunk> c = make_chunks(n, clen);
    b->all_chunks.insert(b->all_chunks.end(), c.begin(), c.end());
    gr.nc = (int64_t)c.size();
    groups.push_back(gr);
    g0 = g1;
  }
  b->chunks.reserve(std::max<size_t>(1, b->all_chunks.size()) * sizeof(BatchChunk));
  HIPC(hipMemcpyAsync(b->probs.p, b->hp.data(), ni * sizeof(DevProblem), hipMemcpyHostToDevice,
                      b->stream));
  HIPC(hipMemcpyAsync(b->ctxs.p, b->hk.data(), ni * sizeof(KCtx), hipMemcpyHostToDevice,
                      b->stream));
  if (!b->all_chunks.empty())
    HIPC(hipMemcpyAsync(b->chunks.p, b->all_chunks.data(),
                        b->all_chunks.size() * sizeof(BatchChunk), hipMemcpyHostToDevice,
                        b->stream));
  // the sources of every entry that starts batches, in one launch
  b->sjobs.clear();
  b->ssrc.clear();
  int64_t stotal = 0;
  int sD = 1, sL = 1;
  for (int j = 0; j < ne; j++) {
    BatchEntry& it = b->items[j];
    if (it.mid_chain) continue;
    abx_problem* p = it.p;
    const int f = first_item[j];
    const int64_t items = (j + 1 < ne ? first_item[j + 1] : ni) - f;  // the entry's batches
    const int64_t total = items * it.n;
    if (total > 0) {
      b->sjobs.push_back(SampleJob{stotal, total, p->run_hist, p->bank_a.as<abx_particle>(),
                                   &ds[f].err, &ds[f].err_hist, pos[f], 0});
      b->ssrc.push_back(p->S);
      stotal += total;
      sD = std::max(sD, p->P.trk_D);
      sL = std::max(sL, p->P.trk_L);
    }
    p->run_hist += (uint64_t)total;
  }
  if (!b->sjobs.empty()) {
    b->sjob_d.reserve(b->sjobs.size() * sizeof(SampleJob));
    b->ssrc_d.reserve(b->ssrc.size() * sizeof(DevSources));
    HIPC(hipMemcpyAsync(b->sjob_d.p, b->sjobs.data(), b->sjobs.size() * sizeof(SampleJob),
                        hipMemcpyHostToDevice, b->stream));
    HIPC(hipMemcpyAsync(b->ssrc_d.p, b->ssrc.data(), b->ssrc.size() * sizeof(DevSources),
                        hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_sample_jobs, dim3((unsigned)blocks_for(stotal, 256)), dim3(256),
                       (size_t)256 * trk_bytes(sD, sL), b->stream, b->probs.as<DevProblem>(),
                       b->ssrc_d.as<DevSources>(), b->sjob_d.as<SampleJob>(),
                       (int)b->sjobs.size(), stotal, sD, sL);
    HIPC(hipGetLastError());
  }
  // one chunk counter per group (groups run concurrently)
  b->next.reserve(std::max<size_t>(8, groups.size()) * sizeof(unsigned long long));
  // the transport: counters zeroed, then one launch per group on the side
  // streams, joined back into the batch stream
  const int used = (int)std::min<size_t>(NS, groups.size());
  auto enqueue = [&]() -> hipError_t {
    hipError_t e;
    const int ng = (int)std::max<size_t>(8, groups.size());
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)blocks_for(ng, 256)), dim3(256), 0, b->stream,
                       b->next.as<unsigned long long>(), ng);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(b->ready, b->stream)) != hipSuccess) return e;
    for (int k = 0; k < used; k++)
      if ((e = hipStreamWaitEvent(b->aux[k], b->ready, 0)) != hipSuccess) return e;
    for (size_t gi = 0; gi < groups.size(); gi++) {
      const Group& gr = groups[gi];
      if (gr.nc > 0 &&
          (e = launch_chunks_variant(gr.key, b->probs.as<DevProblem>() + gr.g0,
                                     b->ctxs.as<KCtx>() + gr.g0, b->chunks.as<BatchChunk>() + gr.c0,
                                     (int)gr.nc, gr.D, gr.L, gr.blob_max,
                                     reinterpret_cast<unsigned*>(b->next.as<unsigned long long>() + gi),
                                     b->block, b->aux[gi % NS])) != hipSuccess)
        return e;
    }
    for (int k = 0; k < used; k++) {
      if ((e = hipEventRecord(b->done[k], b->aux[k])) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(b->stream, b->done[k], 0)) != hipSuccess) return e;
    }
    return hipSuccess;
  };
  // The step's launch signature: what the captured launches fix - the
  // kernel variant, grid, block, LDS and chunk count of every group and the
  // device arrays they read.  The problems, contexts and chunks themselves
  // are uploaded into those arrays before every launch or replay.
  std::vector<int64_t> sig;
  sig.push_back(b->block);
  sig.push_back((int64_t)groups.size());
  for (const Group& gr : groups)
    sig.insert(sig.end(), {(int64_t)gr.key, (int64_t)gr.g0, (int64_t)gr.cnt, (int64_t)gr.D,
                           (int64_t)gr.L, (int64_t)gr.blob_max, gr.c0, gr.nc});
  for (const void* q : {b->probs.p, b->ctxs.p, b->chunks.p, b->next.p, b->sec.p})
    sig.push_back((int64_t)(uintptr_t)q);
  const bool repeat = b->graphs && sig == b->last_sig;
  if (!repeat && b->gexec) {
    HIPC(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  HIPC(hipEventRecord(b->e0, b->stream));
  if (repeat && !b->gexec) {
    hipGraph_t g = nullptr;
    HIPC(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t ce = enqueue();
    const hipError_t ee = hipStreamEndCapture(b->stream, &g);
    if (ce != hipSuccess || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      HIPC(ce != hipSuccess ? ce : ee);
    }
    const hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPC(ie);
    b->captures++;
  }
  if (repeat) {
    HIPC(hipGraphLaunch(b->gexec, b->stream));
    b->replays++;
  } else {
    HIPC(enqueue());
  }
  b->last_sig.swap(sig);
  HIPC(hipEventRecord(b->e1, b->stream));
  b->h_scal.resize(ni);
  HIPC(hipMemcpyAsync(b->h_scal.data(), b->scal.p, (size_t)ni * sizeof(BatchScalars),
                      hipMemcpyDeviceToHost, b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->kernel_ms += ms;
  // 3. per-problem bookkeeping, the concurrent batches of an entry recorded
  // in history order
  std::vector<BatchEntry> keep;
  keep.reserve(ne);
  b->rjobs.clear();
  int64_t rtotal = 0;
  // the tally part of a record: gathered into one launch after the loop
  auto record_tallies = [&](abx_problem* p, double* gen0, int kb, int g0) {
    const int64_t nb = p->T.tally_bins;
    if (nb <= 0) return;
    b->rjobs.push_back(RecordJob{rtotal, nb, gen0, p->tally_more.as<double>(),
                                 p->tally_avg.as<double>(), p->tally_var.as<double>(),
                                 (double)g0, kb, 0});
    rtotal += nb;
  };
  for (int j = 0; j < ne; j++) {
    BatchEntry& it = b->items[j];
    abx_problem* p = it.p;
    const HostTables& T = p->T;
    const int f = first_item[j];
    const int nk = (j + 1 < ne ? first_item[j + 1] : ni) - f;
    if (nk > 1) record_tallies(p, li[f].gen, nk, p->tally_g);
    for (int k = 0; k < nk; k++) {
      const LaunchItem& x = li[f + k];
      const BatchScalars& h = b->h_scal[f + k];
      TransportOut r = collect_scalars(h.sc, h.cn, h.err, h.err_hist);
      add_diag(p->run_diag, r.diag);
      p->run_histories += x.n;
      if (nk > 1) {
        // a whole batch of its own; its tallies are folded in by the
        // record job registered above
        push_gen_values(p, r.scores, (double)T.st.nparticles);
        record_scores(p, r.scores, (double)T.st.nparticles);
        p->gen_scores = abx_scores{};
        it.batches_left--;
        continue;
      }
      add_scores(it.acc, r.scores);
      int64_t n_sites = 0;
      if (T.st.mode == ABX_MODIFIED_FIXED_SOURCE)
        n_sites = order_fission_bank(p, it.n, h.n_stage, h.n_hole);
      if (n_sites > 0) {
        if (p->cancel) cancel_device(p, p->ordered.as<abx_banked>(), n_sites);
        p->bank_b.reserve((size_t)n_sites * sizeof(abx_particle));
        hipLaunchKernelGGL(k_rebank, dim3((unsigned)blocks_for(n_sites, 256)), dim3(256), 0,
                           p->stream, p->ordered.as<abx_banked>(), n_sites, 1.0, p->run_hist,
                           T.st.seed, T.st.stride, 1, 0, p->bank_b.as<abx_particle>());
        HIPC(hipGetLastError());
        p->run_hist += (uint64_t)n_sites;
        std::swap(p->bank_a, p->bank_b);
        p->bank_n = (size_t)n_sites;
        it.n = n_sites;
        it.mid_chain = true;
      } else if (sliceable(p) && (it.done += it.n) < T.st.nparticles) {
        // more slices of this batch follow
      } else {
        it.mid_chain = false;
        it.done = 0;
        record_tallies(p, p->tally_gen.as<double>(), 1, p->tally_g);
        push_gen_values(p, it.acc, (double)T.st.nparticles);
        record_scores(p, it.acc, (double)T.st.nparticles);
        p->gen_scores = abx_scores{};
        it.batches_left--;
      }
      // fission-bank ordering, cancellation and re-banking ran on the
      // problem's own stream; the next step's launches must see them
      if (!sliceable(p)) HIPC(hipStreamSynchronize(p->stream));
    }
    if (it.batches_left > 0) {
      keep.push_back(it);
    } else {
      b->used -= std::min(b->used, it.foot);
      finished.push_back(p);
    }
  }
  b->items.swap(keep);
  if (!b->rjobs.empty()) {
    b->rjob_d.reserve(b->rjobs.size() * sizeof(RecordJob));
    HIPC(hipMemcpyAsync(b->rjob_d.p, b->rjobs.data(), b->rjobs.size() * sizeof(RecordJob),
                        hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_record_jobs, dim3((unsigned)blocks_for(rtotal, 256)), dim3(256), 0,
                       b->stream, b->rjob_d.as<RecordJob>(), (int)b->rjobs.size(), rtotal);
    HIPC(hipGetLastError());
  }
  // the records above ran on the batch stream; finished problems are read
  // on their own streams
  HIPC(hipStreamSynchronize(b->stream));
}

}  // namespace

extern "C" {

int abx_batch_create(size_t budget_bytes, abx_batch** out) {
  ABX_TRY({
    if (!out) fail(ABX_E_INVALID, "null argument");
    auto b = std::make_unique<abx_batch>();
    b->budget = budget_bytes;
    HIPC(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    *out = b.release();
  });
}

void abx_batch_destroy(abx_batch* b) { delete b; }

int abx_problem_footprint(abx_problem* p, size_t* bytes) {
  ABX_TRY({
    if (!p || !bytes) fail(ABX_E_INVALID, "null argument");
    *bytes = problem_footprint(p);
  });
}

int abx_batch_need(abx_batch* b, abx_problem* p, int nbatches, size_t* bytes) {
  ABX_TRY({
    if (!b || !p || !bytes || nbatches < 1) fail(ABX_E_INVALID, "batch arguments");
    *bytes = batch_need(b, p, nbatches);
  });
}

int abx_batch_add(abx_batch* b, abx_problem* p, int nbatches) {
  ABX_TRY({
    if (!b || !p || nbatches < 1) fail(ABX_E_INVALID, "batch arguments");
    if (p->T.st.mode == ABX_K_EIGENVALUE)
      fail(ABX_E_INVALID, "k-eigenvalue problems run with abx_run, not in a batch");
    for (const auto& it : b->items)
      if (it.p == p) fail(ABX_E_INVALID, "problem already attached");
    const size_t need = batch_need(b, p, nbatches);
    if (b->budget && b->used + need > b->budget)
      fail(ABX_E_BUDGET, "problem does not fit the batch's device budget");
    run_reset(p);
    // run_reset clears the tally arrays on the problem's own stream; the
    // batch launches on its streams, which are not ordered after it
    HIPC(hipStreamSynchronize(p->stream));
    p->run_diag = abx_diag{};
    p->run_histories = 0;
    b->used += need;
    b->items.push_back(BatchEntry{p, nbatches, false, 0, abx_scores{}, 0, need});
  });
}

int abx_batch_step(abx_batch* b, abx_problem** finished, int cap, int* n_finished,
                   int* n_active) {
  ABX_TRY({
    if (!b) fail(ABX_E_INVALID, "null argument");
    std::vector<abx_problem*> done;
    batch_step(b, done);
    if (n_finished) *n_finished = (int)done.size();
    if ((int)done.size() > cap && finished) fail(ABX_E_CAPACITY, "finished list too small");
    for (size_t k = 0; k < done.size() && finished; k++) finished[k] = done[k];
    if (n_active) *n_active = (int)b->items.size();
  });
}

int abx_batch_set_step_histories(abx_batch* b, int64_t histories) {
  ABX_TRY({
    if (!b || histories < 0) fail(ABX_E_INVALID, "step histories");
    if (!b->items.empty()) fail(ABX_E_INVALID, "set the step cap before attaching problems");
    b->step_hist = histories;
  });
}

int abx_batch_set_params(abx_batch* b, int64_t chunk, int64_t multi_hist) {
  ABX_TRY({
    if (!b || multi_hist < 0) fail(ABX_E_INVALID, "batch parameters");
    // attached problems were admitted (budget) with the concurrency these
    // parameters decide
    if (!b->items.empty()) fail(ABX_E_INVALID, "set batch parameters before attaching problems");
    b->chunk = chunk;
    b->multi_hist = multi_hist > 0 ? multi_hist : (int64_t(1) << 23);
  });
}

int abx_batch_set_block_size(abx_batch* b, int block_size) {
  ABX_TRY({
    if (!b) fail(ABX_E_INVALID, "null argument");
    if (block_size != 64 && block_size != 128 && block_size != 256)
      fail(ABX_E_INVALID, "block_size must be 64, 128 or 256");
    b->block = block_size;
  });
}

int abx_batch_set_graphs(abx_batch* b, int on) {
  ABX_TRY({
    if (!b) fail(ABX_E_INVALID, "null argument");
    b->graphs = on != 0;
  });
}

int abx_batch_graph_stats(abx_batch* b, int64_t* captures, int64_t* replays) {
  ABX_TRY({
    if (!b || !captures || !replays) fail(ABX_E_INVALID, "null argument");
    *captures = b->captures;
    *replays = b->replays;
  });
}

int abx_batch_kernel_seconds(abx_batch* b, double* seconds) {
  ABX_TRY({
    if (!b || !seconds) fail(ABX_E_INVALID, "null argument");
    *seconds = b->kernel_ms * 1e-3;
  });
}

int abx_problem_stats(abx_problem* p, abx_run_stats* stats) {
  ABX_TRY({
    if (!p || !stats) fail(ABX_E_INVALID, "null argument");
    *stats = abx_run_stats{};
    fill_statistics(p, stats);
    stats->diag = p->run_diag;
    stats->histories = p->run_histories;
  });
}

}  // extern "C"

// ------------------------------------------------ transport without allocation
// abx_transport_noalloc keeps every device buffer of a transport call beyond
// the caller's bank and output (scores, counters, the history counter,
// fission staging, per-parent counts and offsets, the secondary stack) in
// one file-scope device array.  Each GPU holds its own copy at its own
// address, taken once per device with hipGetSymbolAddress; host code only
// ever passes that address, never the array's name.
namespace {

constexpr size_t kNoAllocBytes = (size_t)64 << 20;
constexpr int64_t kNoAllocMaxHistories = ABX_NOALLOC_MAX_HISTORIES;
static_assert(kNoAllocBytes == ABX_NOALLOC_SCRATCH_BYTES, "scratch size");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (taken with that device
// current, then restored).
unsigned char* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) fail(ABX_E_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPC(hipGetDevice(&cur));
    if (cur != dev) HIPC(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIPC(hipSetDevice(cur));
    hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) fail(ABX_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

// Slices of the scratch array for one call over n histories.
struct NoAllocLayout {
  size_t scalars, daughters, offsets, block_sums, stage, stage_parent, stage_daughter, sec, end;
};

constexpr size_t kScalarBytes = 256;  // scores[8] counters[8] err err_hist n_stage[2] next total

NoAllocLayout noalloc_layout(int64_t n, int64_t stage_cap, int sec_cap) {
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const int64_t nb = (n + 1023) / 1024;
  NoAllocLayout L{};
  size_t o = 0;
  L.scalars = o;
  o += up(kScalarBytes);
  L.daughters = o;
  o += up((size_t)n * sizeof(uint32_t));
  L.offsets = o;
  o += up((size_t)n * sizeof(uint64_t));
  L.block_sums = o;
  o += up((size_t)(nb + 1) * sizeof(uint64_t));
  L.stage = o;
  o += up((size_t)stage_cap * sizeof(abx_banked));
  L.stage_parent = o;
  o += up((size_t)stage_cap * sizeof(uint32_t));
  L.stage_daughter = o;
  o += up((size_t)stage_cap * sizeof(uint32_t));
  L.sec = o;
  o += up((size_t)(sec_cap ? n + 256 : 0) * sec_cap * sizeof(SecRec));
  L.end = o;
  return L;
}

// The worst case (a banking problem with the secondary stack) at the
// largest accepted call fits the array.
static_assert(kNoAllocMaxHistories * (int64_t)(sizeof(uint32_t) + sizeof(uint64_t)) +
                      (4 * kNoAllocMaxHistories + 4096) *
                          (int64_t)(sizeof(abx_banked) + 2 * sizeof(uint32_t)) +
                      (kNoAllocMaxHistories + 256) * 16 * (int64_t)sizeof(SecRec) + (1 << 20) <=
                  (int64_t)kNoAllocBytes,
              "scratch too small for the largest call");

}  // namespace

extern "C" int abx_transport_noalloc(abx_problem* p, const abx_particle* d_bank, size_t n,
                                     double kcol, int converged, abx_banked* d_fission_out,
                                     size_t cap, size_t* n_out, abx_scores* scores,
                                     abx_diag* diag) {
  ABX_TRY({
    if (!p || (!d_bank && n)) fail(ABX_E_INVALID, "null argument");
    if ((int64_t)n > kNoAllocMaxHistories)
      fail(ABX_E_INVALID, "abx_transport_noalloc takes at most " +
                              std::to_string(kNoAllocMaxHistories) + " histories per call, got " +
                              std::to_string(n));
    const HostTables& T = p->T;
    const int64_t nn = (int64_t)n;
    const bool banks = T.st.mode != ABX_FIXED_SOURCE;
    const int sec_cap = needs_secondaries(T) ? 16 : 0;
    const int64_t stage_cap = banks ? 4 * nn + 4096 : 0;
    const NoAllocLayout L = noalloc_layout(nn, stage_cap, sec_cap);
    if (L.end > kNoAllocBytes) fail(ABX_E_INVALID, "scratch layout exceeds the scratch array");
    unsigned char* base = scratch_on(p->device);
    double* d_scores = reinterpret_cast<double*>(base + L.scalars);
    unsigned long long* d_counters = reinterpret_cast<unsigned long long*>(base + L.scalars + 64);
    int32_t* d_err = reinterpret_cast<int32_t*>(base + L.scalars + 128);
    unsigned long long* d_err_hist = reinterpret_cast<unsigned long long*>(base + L.scalars + 136);
    unsigned long long* d_nst = reinterpret_cast<unsigned long long*>(base + L.scalars + 144);
    unsigned long long* d_next = reinterpret_cast<unsigned long long*>(base + L.scalars + 160);
    unsigned long long* d_total = reinterpret_cast<unsigned long long*>(base + L.scalars + 168);
    uint32_t* d_daughters = reinterpret_cast<uint32_t*>(base + L.daughters);
    uint64_t* d_offsets = reinterpret_cast<uint64_t*>(base + L.offsets);
    uint64_t* d_bsums = reinterpret_cast<uint64_t*>(base + L.block_sums);
    // the previous call's scalars are still there: clear them with a kernel
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, p->stream, d_scores,
                       (int64_t)(kScalarBytes / sizeof(double)), 0.0);
    HIPC(hipGetLastError());
    p->last_sites = 0;
    KCtx K{};
    K.in = d_bank;
    K.n = nn;
    K.kcol = kcol;
    K.converged = converged;
    K.sec_cap = sec_cap;
    K.sec = sec_cap ? reinterpret_cast<SecRec*>(base + L.sec) : nullptr;
    K.stage = banks ? reinterpret_cast<abx_banked*>(base + L.stage) : nullptr;
    K.stage_parent = banks ? reinterpret_cast<uint32_t*>(base + L.stage_parent) : nullptr;
    K.stage_daughter = banks ? reinterpret_cast<uint32_t*>(base + L.stage_daughter) : nullptr;
    K.n_stage = d_nst;
    K.n_hole = d_nst + 1;
    K.stage_cap = stage_cap;
    K.daughters = banks ? d_daughters : nullptr;
    K.scores = d_scores;
    K.counters = d_counters;
    K.err = d_err;
    K.err_hist = d_err_hist;
    if (nn > 0)
      HIPC(history_transport(p->P, p->T.extended, K, p->stream, sec_cap > 0, d_next,
                             p->tuning));
    double sc[8];
    unsigned long long cn[8], nst[2] = {0, 0};
    int32_t err = 0;
    unsigned long long err_hist = 0;
    HIPC(hipMemcpyAsync(sc, d_scores, sizeof(sc), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipMemcpyAsync(cn, d_counters, sizeof(cn), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipMemcpyAsync(&err, d_err, sizeof(err), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipMemcpyAsync(&err_hist, d_err_hist, sizeof(err_hist), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipMemcpyAsync(nst, d_nst, sizeof(nst), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    TransportOut r = collect_scalars(sc, cn, err, err_hist);
    add_scores(p->gen_scores, r.scores);
    if (scores) *scores = r.scores;
    if (diag) *diag = r.diag;
    if (n_out) *n_out = 0;
    if (!banks || nn == 0) return ABX_OK;
    // order the staged sites (parent order, then push order) into the
    // caller's buffer: scan of the per-parent counts, all on the device
    if ((int64_t)nst[0] > stage_cap) fail(ABX_E_CAPACITY, "fission staging overflow");
    const int64_t nb = (nn + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(1024), 0, p->stream, d_daughters,
                       d_offsets, d_bsums, nn);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, p->stream, d_bsums, nb, d_total);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_add_block_offsets, dim3((unsigned)nb), dim3(1024), 0, p->stream,
                       d_offsets, d_bsums, nn);
    HIPC(hipGetLastError());
    unsigned long long total = 0;
    HIPC(hipMemcpyAsync(&total, d_total, sizeof(total), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    if (nst[1] > nst[0] || total != nst[0] - nst[1]) fail(ABX_E_INVALID, "fission site count mismatch");
    if (n_out) *n_out = (size_t)total;
    if (total == 0) return ABX_OK;
    if (!d_fission_out || total > cap)
      fail(ABX_E_CAPACITY, "fission_out holds " + std::to_string(cap) + " sites, " +
                               std::to_string(total) + " were banked");
    hipLaunchKernelGGL(k_order_bank, dim3((unsigned)blocks_for((int64_t)nst[0], 256)), dim3(256), 0,
                       p->stream, K.stage, K.stage_parent, K.stage_daughter, d_offsets,
                       (int64_t)nst[0], nn, (uint64_t)total, d_fission_out, d_err,
                       (const abx_parent_info*)nullptr, (abx_parent_info*)nullptr);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(&err, d_err, sizeof(err), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    if (err) fail(err, "fission bank ordering: a staging slot names no parent of this call");
  });
}
