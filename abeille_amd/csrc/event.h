// event.h - event-based transport engine.
//
// Instead of one lane carrying a history from birth to death (the
// reference's per-history OpenMP loop, src/delta_tracker.cpp:83-229), the
// bank lives in HBM as structure-of-arrays and each kernel advances every
// queued particle by ONE event:
//   k_ev_init     locate the source particle (Tracker construction)
//   k_ev_flight   sample a flight, move / relocate, boundary conditions,
//                 track-length tally, real/virtual decision
//   k_ev_collide  collision physics (tallies, k scores, fission, implicit
//                 capture, roulette, scatter)
// and appends the particle's slot to the queue of its next event with one
// wavefront-aggregated atomic (ballot + popcount).  Queues shrink as
// histories end; when few particles remain (the long tail) the rest are
// finished by k_ev_tail, which runs the whole remaining history in a lane.
// Per-particle arithmetic is the same physics.h code as the history kernel,
// so results are bit-identical per history.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <atomic>
#include <vector>

#include "kernels.h"

#ifndef ABX_KERNELS_ONLY
#include <rocprim/device/device_radix_sort.hpp>
#endif

// Structure-of-arrays particle + cursor state, one slot per history.
struct EvState {
  double *rx, *ry, *rz, *ux, *uy, *uz, *E, *w, *w2, *bx, *by, *bz;
  uint64_t *rng, *hid, *fam;
  int32_t *calls, *daughters, *secid, *g, *m, *nsec, *nrec;
  int32_t *np, *cell, *token;
  int32_t* pad;   // [D][n]
  int32_t* tile;  // [3 D][n]
  double* lv;     // [3 L][n]
  int64_t n;
  int32_t D, L;
};

__device__ __forceinline__ void ev_load(const EvState& S, int64_t i, Part& p, TrkT<true>& t, int& m) {
  p.r = V3{S.rx[i], S.ry[i], S.rz[i]};
  p.u = V3{S.ux[i], S.uy[i], S.uz[i]};
  p.rb = V3{S.bx[i], S.by[i], S.bz[i]};
  p.E = S.E[i];
  p.w = S.w[i];
  p.w2 = S.w2[i];
  p.rng = S.rng[i];
  p.hid = S.hid[i];
  p.fam = S.fam[i];
  p.calls = (uint32_t)S.calls[i];
  p.daughters = (uint32_t)S.daughters[i];
  p.sec_id = S.secid[i];
  p.g = S.g[i];
  p.alive = 1;
  m = S.m[i];
  t.np = S.np[i];
  t.cell = S.cell[i];
  t.token = S.token[i];
  t.err = 0;
  const int64_t n = S.n;
  int maxl = 0;
  for (int k = 0; k < t.np; k++) {
    const int32_t w = S.pad[k * n + i];
    t.pad_[k * t.st] = w;
    if (pad_type(w) == PAD_L) {
      t.tile_[(3 * k) * t.st] = S.tile[(3 * k) * n + i];
      t.tile_[(3 * k + 1) * t.st] = S.tile[(3 * k + 1) * n + i];
      t.tile_[(3 * k + 2) * t.st] = S.tile[(3 * k + 2) * n + i];
    }
    const int l = pad_lvl(w);
    maxl = l > maxl ? l : maxl;
  }
  t.set_lv(0, p.r);  // frame 0 is the global position
  for (int l = 1; l <= maxl; l++)
    t.set_lv(l, V3{S.lv[(3 * l) * n + i], S.lv[(3 * l + 1) * n + i], S.lv[(3 * l + 2) * n + i]});
}

__device__ __forceinline__ void ev_store(const EvState& S, int64_t i, const Part& p, const TrkT<true>& t,
                                         int m) {
  S.rx[i] = p.r.x;
  S.ry[i] = p.r.y;
  S.rz[i] = p.r.z;
  S.ux[i] = p.u.x;
  S.uy[i] = p.u.y;
  S.uz[i] = p.u.z;
  S.E[i] = p.E;
  S.w[i] = p.w;
  S.w2[i] = p.w2;
  S.rng[i] = p.rng;
  S.calls[i] = (int32_t)p.calls;
  S.daughters[i] = (int32_t)p.daughters;
  S.secid[i] = p.sec_id;
  S.g[i] = p.g;
  S.m[i] = m;
  S.np[i] = t.np;
  S.cell[i] = t.cell;
  S.token[i] = t.token;
  const int64_t n = S.n;
  int maxl = 0;
  for (int k = 0; k < t.np; k++) {
    const int32_t w = t.pad(k);
    S.pad[k * n + i] = w;
    if (pad_type(w) == PAD_L) {
      S.tile[(3 * k) * n + i] = t.tile(k, 0);
      S.tile[(3 * k + 1) * n + i] = t.tile(k, 1);
      S.tile[(3 * k + 2) * n + i] = t.tile(k, 2);
    }
    const int l = pad_lvl(w);
    maxl = l > maxl ? l : maxl;
  }
  for (int l = 1; l <= maxl; l++) {
    const V3 v = t.lv(l);
    S.lv[(3 * l) * n + i] = v.x;
    S.lv[(3 * l + 1) * n + i] = v.y;
    S.lv[(3 * l + 2) * n + i] = v.z;
  }
}

// Append `slot` to a queue with one atomic per wavefront.
__device__ __forceinline__ void wave_push(bool pred, int32_t slot, int32_t* q, unsigned* count) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) {
    const unsigned long long below = mask & ((1ull << lane) - 1ull);
    q[base + (unsigned)__popcll(below)] = slot;
  }
}

// The slot's history is over: publish its per-history outputs.
template <class Sink>
__device__ __forceinline__ void ev_finalize(const KCtx& K, const Sink& s, int64_t i,
                                            const Part& p) {
  if (K.daughters) K.daughters[i] = p.daughters;
  if (K.n_rec) K.n_rec[i] = s.n_rec;
}

template <class Sink>
__device__ __forceinline__ void ev_resume(Sink& s, const EvState& S, int64_t i) {
  s.slot = i;
  s.sslot = i;
  s.nsec = S.nsec[i];
  s.n_rec = S.nrec[i];
}

template <class Sink>
__device__ __forceinline__ void ev_save(const Sink& s, const EvState& S, int64_t i) {
  S.nsec[i] = s.nsec;
  S.nrec[i] = s.n_rec;
}

template <bool TRACE, bool SEC>
__global__ void __launch_bounds__(256) k_ev_init(DevProblem P, KCtx K, EvState S, int32_t* qf,
                                                 unsigned* nqf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Scores6 sc{0, 0, 0, 0, 0, 0};
  GpuSink<TRACE, SEC> s;
  s.init(&K, &P, i);
  bool live = false;
  if (i < K.n) {
    Part p;
    load_part(P, K.in[i], p);
    TrkT<true> t;
    bind_lds_cursor(t, P);
    int m = -1;
    live = start_history(P, s, p, t, m);
    s.c[0] = 1;
    S.bx[i] = p.rb.x;
    S.by[i] = p.rb.y;
    S.bz[i] = p.rb.z;
    S.hid[i] = p.hid;
    S.fam[i] = p.fam;
    S.nsec[i] = 0;
    S.nrec[i] = 0;
    if (live)
      ev_store(S, i, p, t, m);
    else
      ev_finalize(K, s, i, p);
  }
  wave_push(live, (int32_t)i, qf, nqf);
  flush_scores(K, sc, s.c);
}

template <int TRK, bool TRACE, bool SEC, bool SPLIT>
__global__ void ABX_TKERNEL k_ev_flight(DevProblem P, KCtx K, EvState S,
                                                   const int32_t* qin, const unsigned* nqin,
                                                   int32_t* qf, unsigned* nqf, int32_t* qc,
                                                   unsigned* nqc) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Scores6 sc{0, 0, 0, 0, 0, 0};
  GpuSink<TRACE, SEC, SPLIT> s;
  s.init(&K, &P, 0);
  bool to_f = false, to_c = false;
  int32_t slot = 0;
  if (k < (int64_t)*nqin) {
    slot = qin[k];
    ev_resume(s, S, slot);
    Part p;
    TrkT<true> t;
    bind_lds_cursor(t, P);
    int m;
    ev_load(S, slot, p, t, m);
    int ev = 0;
    const bool hc = step_flight<TRK>(P, s, p, t, m, K.converged != 0, sc, ev);
    if (p.alive && hc) {
      to_c = true;
    } else if (step_finish<TRK>(P, s, p, t, m, ev)) {
      to_f = true;
    }
    if (to_f || to_c) {
      ev_store(S, slot, p, t, m);
      ev_save(s, S, slot);
    } else {
      ev_finalize(K, s, slot, p);
    }
  }
  wave_push(to_f, slot, qf, nqf);
  wave_push(to_c, slot, qc, nqc);
  flush_scores(K, sc, s.c);
}

template <int TRK, bool TRACE, bool SEC, bool SPLIT>
__global__ void ABX_TKERNEL k_ev_collide(DevProblem P, KCtx K, EvState S,
                                                    const int32_t* qin, const unsigned* nqin,
                                                    int32_t* qf, unsigned* nqf) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Scores6 sc{0, 0, 0, 0, 0, 0};
  GpuSink<TRACE, SEC, SPLIT> s;
  s.init(&K, &P, 0);
  bool to_f = false;
  int32_t slot = 0;
  if (k < (int64_t)*nqin) {
    slot = qin[k];
    ev_resume(s, S, slot);
    Part p;
    TrkT<true> t;
    bind_lds_cursor(t, P);
    int m;
    ev_load(S, slot, p, t, m);
    int ev = 0;
    step_collision(P, s, p, m, K.kcol, K.converged != 0, sc, ev);
    to_f = step_finish<TRK>(P, s, p, t, m, ev);
    if (to_f) {
      ev_store(S, slot, p, t, m);
      ev_save(s, S, slot);
    } else {
      ev_finalize(K, s, slot, p);
    }
  }
  wave_push(to_f, slot, qf, nqf);
  flush_scores(K, sc, s.c);
}

// Finish the remaining histories in place (flight queue first, then the
// collision queue whose particles have a collision pending).
template <int TRK, bool TRACE, bool SEC, bool SPLIT>
__global__ void ABX_TKERNEL k_ev_tail(DevProblem P, KCtx K, EvState S,
                                                 const int32_t* qf, const unsigned* nqf,
                                                 const int32_t* qc, const unsigned* nqc) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Scores6 sc{0, 0, 0, 0, 0, 0};
  GpuSink<TRACE, SEC, SPLIT> s;
  s.init(&K, &P, 0);
  const int64_t nf = *nqf, nc = *nqc;
  if (k < nf + nc) {
    const bool pending = k >= nf;
    const int32_t slot = pending ? qc[k - nf] : qf[k];
    ev_resume(s, S, slot);
    Part p;
    TrkT<true> t;
    bind_lds_cursor(t, P);
    int m;
    ev_load(S, slot, p, t, m);
    run_loop<TRK>(P, s, p, t, m, K.kcol, K.converged != 0, sc, pending);
    ev_finalize(K, s, slot, p);
  }
  flush_scores(K, sc, s.c);
}

#ifndef ABX_KERNELS_ONLY
// Sort keys of the collision queue: (material, group) of each queued slot,
// so that a sorted queue hands each wavefront particles of one material
// and group (the same table rows, the same fissile / non-fissile branch).
__global__ void __launch_bounds__(256) k_ev_sort_keys(EvState S, const int32_t* q, int64_t n,
                                                      uint16_t* keys) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t slot = q[k];
  keys[k] = (uint16_t)(((S.m[slot] & 0xff) << 8) | (S.g[slot] & 0xff));
}
#endif  // ABX_KERNELS_ONLY

// The transport kernels' variants are compiled in instances.hip.
#include "instances.h"

// Host-side launch code follows; ABX_KERNELS_ONLY leaves it out (a
// translation unit that only instantiates chosen kernels).
#ifndef ABX_KERNELS_ONLY

// Device buffers of the engine (grow-only).
struct EventBuffers {
  void* state = nullptr;
  size_t state_bytes = 0;
  int32_t* q[3] = {nullptr, nullptr, nullptr};
  size_t q_cap = 0;
  unsigned* counts = nullptr;     // device [8]
  unsigned* h_counts = nullptr;   // pinned host mirror
  std::vector<hipEvent_t> evs;
  int64_t tail_threshold = 0;     // 0 = automatic
  int32_t sort_every = 0;         // sort the collision queue every n-th iteration (0 = never)
  uint16_t* keys[2] = {nullptr, nullptr};
  int32_t* q_sorted = nullptr;
  void* sort_temp = nullptr;
  size_t sort_temp_bytes = 0;
  int64_t iterations = 0;
  int64_t flight_items = 0;       // particles processed by k_ev_flight (last call)
  void release() {
    if (state) (void)hipFree(state);
    for (auto& x : q)
      if (x) (void)hipFree(x);
    if (counts) (void)hipFree(counts);
    if (h_counts) (void)hipHostFree(h_counts);
    for (auto& x : keys)
      if (x) (void)hipFree(x);
    if (q_sorted) (void)hipFree(q_sorted);
    if (sort_temp) (void)hipFree(sort_temp);
    keys[0] = keys[1] = nullptr;
    q_sorted = nullptr;
    sort_temp = nullptr;
    sort_temp_bytes = 0;
    for (auto e : evs) (void)hipEventDestroy(e);
    evs.clear();
    state = nullptr;
    counts = nullptr;
    h_counts = nullptr;
    for (auto& x : q) x = nullptr;
  }
};

static inline size_t ev_state_bytes(int64_t n, int D, int L) {
  return (size_t)n * (12 * sizeof(double) + 3 * sizeof(uint64_t) + 10 * sizeof(int32_t) +
                      (size_t)D * 4 * sizeof(int32_t) + (size_t)L * 3 * sizeof(double)) +
         40 * 256;
}

static inline EvState ev_carve(void* base, int64_t n, int D, int L) {
  EvState S;
  char* b = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* r = b;
    b += (bytes + 255) & ~(size_t)255;
    return r;
  };
  double** d[] = {&S.rx, &S.ry, &S.rz, &S.ux, &S.uy, &S.uz, &S.E, &S.w, &S.w2, &S.bx, &S.by, &S.bz};
  for (auto pp : d) *pp = reinterpret_cast<double*>(take(n * sizeof(double)));
  uint64_t** u[] = {&S.rng, &S.hid, &S.fam};
  for (auto pp : u) *pp = reinterpret_cast<uint64_t*>(take(n * sizeof(uint64_t)));
  int32_t** ii[] = {&S.calls, &S.daughters, &S.secid, &S.g, &S.m, &S.nsec, &S.nrec, &S.np, &S.cell, &S.token};
  for (auto pp : ii) *pp = reinterpret_cast<int32_t*>(take(n * sizeof(int32_t)));
  S.pad = reinterpret_cast<int32_t*>(take((size_t)D * n * sizeof(int32_t)));
  S.tile = reinterpret_cast<int32_t*>(take((size_t)3 * D * n * sizeof(int32_t)));
  S.lv = reinterpret_cast<double*>(take((size_t)3 * L * n * sizeof(double)));
  S.n = n;
  S.D = D;
  S.L = L;
  return S;
}

// Runs one transport() with the event engine.  `kernel_ms` receives the
// summed duration of the flight kernels (the dominant kernel), measured with
// HIP events on `s`, and `launches` their count, when `timing` is set.
template <int TRK, bool TRACE, bool SEC, bool SPLIT>
static hipError_t event_transport_t(const DevProblem& P, const KCtx& K, EventBuffers& B,
                                    hipStream_t s, bool timing, double* kernel_ms,
                                    int64_t* launches) {
  hipError_t e;
  const int64_t n = K.n;
  const int D = P.trk_D, L = P.trk_L;
  const int bs = 256;
  const size_t lds = (size_t)bs * trk_bytes(D, L);
  const size_t need = ev_state_bytes(n, D, L) + 64 * 256;
  if (need > B.state_bytes) {
    if (B.state) (void)hipFree(B.state);
    B.state = nullptr;
    if ((e = hipMalloc(&B.state, need)) != hipSuccess) return e;
    B.state_bytes = need;
  }
  if ((size_t)n > B.q_cap) {
    for (auto& x : B.q) {
      if (x) (void)hipFree(x);
      x = nullptr;
      if ((e = hipMalloc(&x, (size_t)n * sizeof(int32_t))) != hipSuccess) return e;
    }
    for (auto& x : B.keys) {
      if (x) (void)hipFree(x);
      x = nullptr;
    }
    if (B.q_sorted) (void)hipFree(B.q_sorted);
    B.q_sorted = nullptr;
    B.q_cap = (size_t)n;
  }
  if (B.sort_every > 0 && !B.q_sorted) {
    for (auto& x : B.keys)
      if ((e = hipMalloc(&x, B.q_cap * sizeof(uint16_t))) != hipSuccess) return e;
    if ((e = hipMalloc(&B.q_sorted, B.q_cap * sizeof(int32_t))) != hipSuccess) return e;
    size_t tb = 0;
    if ((e = rocprim::radix_sort_pairs(nullptr, tb, B.keys[0], B.keys[1], B.q[2], B.q_sorted,
                                       B.q_cap, 0, 16, s)) != hipSuccess)
      return e;
    if (tb > B.sort_temp_bytes) {
      if (B.sort_temp) (void)hipFree(B.sort_temp);
      B.sort_temp = nullptr;
      if ((e = hipMalloc(&B.sort_temp, tb)) != hipSuccess) return e;
      B.sort_temp_bytes = tb;
    }
  }
  if (!B.counts) {
    if ((e = hipMalloc(&B.counts, 8 * sizeof(unsigned))) != hipSuccess) return e;
    if ((e = hipHostMalloc(&B.h_counts, 8 * sizeof(unsigned), hipHostMallocDefault)) != hipSuccess)
      return e;
  }
  const EvState S = ev_carve(B.state, n, D, L);
  // counts: [0] flight queue A, [1] flight queue B, [2] collision queue
  int32_t* qa = B.q[0];
  int32_t* qb = B.q[1];
  int32_t* qc = B.q[2];
  unsigned* ca = B.counts + 0;
  unsigned* cb = B.counts + 1;
  unsigned* cc = B.counts + 2;
  if ((e = hipMemsetAsync(B.counts, 0, 8 * sizeof(unsigned), s)) != hipSuccess) return e;
  hipLaunchKernelGGL((k_ev_init<TRACE, SEC>), dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), lds, s,
                     P, K, S, qa, ca);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t tail = B.tail_threshold > 0 ? B.tail_threshold : std::max<int64_t>(8192, n / 64);
  size_t ev_used = 0;
  auto get_event = [&]() -> hipEvent_t {
    if (ev_used == B.evs.size()) {
      hipEvent_t x;
      if (hipEventCreate(&x) != hipSuccess) return nullptr;
      B.evs.push_back(x);
    }
    return B.evs[ev_used++];
  };
  if ((e = hipMemcpyAsync(B.h_counts, B.counts, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  int64_t nf = B.h_counts[0], nc = 0;
  int64_t iters = 0, items = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timed;
  while (nf + nc > 0) {
    if (nf + nc <= tail) {
      hipLaunchKernelGGL((k_ev_tail<TRK, TRACE, SEC, SPLIT>), dim3((unsigned)((nf + nc + bs - 1) / bs)),
                         dim3(bs), lds, s, P, K, S, qa, ca, qc, cc);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      break;
    }
    // collisions first (they refill the flight queue), then flights
    if ((e = hipMemsetAsync(cb, 0, sizeof(unsigned), s)) != hipSuccess) return e;
    if (nc > 0) {
      // the collision queue, sorted by (material, group) on the chosen cadence
      const int32_t* qcol = qc;
      if (B.sort_every > 0 && iters % B.sort_every == 0 && nc > 1) {
        hipLaunchKernelGGL(k_ev_sort_keys, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, S,
                           qc, nc, B.keys[0]);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        size_t tb = B.sort_temp_bytes;
        if ((e = rocprim::radix_sort_pairs(B.sort_temp, tb, B.keys[0], B.keys[1], qc, B.q_sorted,
                                           (size_t)nc, 0, 16, s)) != hipSuccess)
          return e;
        qcol = B.q_sorted;
      }
      hipLaunchKernelGGL((k_ev_collide<TRK, TRACE, SEC, SPLIT>), dim3((unsigned)((nc + bs - 1) / bs)),
                         dim3(bs), lds, s, P, K, S, qcol, cc, qa, ca);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int64_t nf_all = nf + nc;  // upper bound on the flight queue
    if ((e = hipMemsetAsync(cc, 0, sizeof(unsigned), s)) != hipSuccess) return e;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing) {
      t0 = get_event();
      t1 = get_event();
      (void)hipEventRecord(t0, s);
    }
    hipLaunchKernelGGL((k_ev_flight<TRK, TRACE, SEC, SPLIT>), dim3((unsigned)((nf_all + bs - 1) / bs)),
                       dim3(bs), lds, s, P, K, S, qa, ca, qb, cb, qc, cc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (timing) {
      (void)hipEventRecord(t1, s);
      timed.push_back({t0, t1});
    }
    if ((e = hipMemcpyAsync(B.h_counts, B.counts, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    items += B.h_counts[ca - B.counts];  // input of this flight launch
    std::swap(qa, qb);
    std::swap(ca, cb);
    nf = B.h_counts[ca - B.counts];
    nc = B.h_counts[2];
    iters++;
  }
  B.iterations = iters;
  B.flight_items = items;
  if (timing) {
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    double ms = 0.0;
    for (auto& pr : timed) {
      float x = 0.f;
      (void)hipEventElapsedTime(&x, pr.first, pr.second);
      ms += x;
    }
    *kernel_ms = ms;
    *launches = (int64_t)timed.size();
  }
  return hipSuccess;
}

// Resident blocks of a transport kernel on this device (persistent grid).
template <class Kern>
static int64_t resident_blocks(Kern kernel, int bs, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bs, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (int64_t)cus * per_cu;
}

// LDS one block may allocate on the current device (160 KB on gfx950).
static size_t lds_block_limit() {
  static std::atomic<size_t> cached[64];  // per device; batches may step from several threads
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  size_t c = cached[dev].load(std::memory_order_relaxed);
  if (!c) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
      v = 65536;
    c = (size_t)v;
    cached[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

// Whether a kernel's LDS-staged variant is worth launching: the tables fit
// next to the cursors, and the extra LDS costs no resident blocks (the
// staged variant gains ~10 % from LDS reads, losing a block per CU costs
// more).
template <class KS, class KU>
static bool stage_pays(KS staged, KU unstaged, int bs, size_t cur, size_t blob) {
  if (cur + blob > lds_block_limit()) return false;
  return resident_blocks(staged, bs, cur + blob) >= resident_blocks(unstaged, bs, cur);
}

// `next` is a device counter (zeroed here) from which lanes take histories.
template <int TRK, bool TRACE, bool SEC, bool SPLIT, bool EXT>
static hipError_t history_transport_t(const DevProblem& P, const KCtx& K, hipStream_t s,
                                      unsigned long long* next, const abx_tuning& tu) {
  const int bs = tu.block_size;
  // Stage the tables in LDS when they fit next to the cursors.
  const size_t cur = cursor_lds_bytes(bs, P.trk_D, P.trk_L);
  const bool stage = tu.stage_tables &&
                     stage_pays(k_history<TRK, TRACE, SEC, SPLIT, true, false, EXT>,
                                k_history<TRK, TRACE, SEC, SPLIT, false, false, EXT>, bs, cur,
                                (size_t)P.blob_bytes);
  const size_t lds = stage ? cur + (size_t)P.blob_bytes : cur;
  hipError_t e;
  if ((e = hipMemsetAsync(next, 0, sizeof(unsigned long long), s)) != hipSuccess) return e;
  const int64_t want = (K.n + bs - 1) / bs;
  int64_t grid = stage ? resident_blocks(k_history<TRK, TRACE, SEC, SPLIT, true, false, EXT>, bs, lds)
                       : resident_blocks(k_history<TRK, TRACE, SEC, SPLIT, false, false, EXT>, bs, lds);
  if (tu.blocks_per_cu > 0) {
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
      return e;
    grid = std::min<int64_t>(grid, (int64_t)cus * tu.blocks_per_cu);
  }
  const int64_t blocks = std::min<int64_t>(want, grid);
  if (stage)
    hipLaunchKernelGGL((k_history<TRK, TRACE, SEC, SPLIT, true, false, EXT>), dim3((unsigned)blocks), dim3(bs),
                       lds, s, P, K, next);
  else
    hipLaunchKernelGGL((k_history<TRK, TRACE, SEC, SPLIT, false, false, EXT>), dim3((unsigned)blocks), dim3(bs),
                       lds, s, P, K, next);
  return hipGetLastError();
}

// Runtime dispatch over the compiled variants (tracking mode x trace x
// secondary stack x precision).  Carter tracking always carries the stack
// (splitting); traces are recorded in FP64 mode only.
template <int V>
using IntC = std::integral_constant<int, V>;
template <bool V>
using BoolC = std::integral_constant<bool, V>;

template <class Fn>
static hipError_t dispatch_variant_t(int trk, bool trace, bool sec, bool split, Fn&& fn) {
  if (split) {
    if (trace) return hipErrorInvalidValue;
    if (trk == 0)
      return sec ? fn(IntC<0>{}, BoolC<false>{}, BoolC<true>{}, BoolC<true>{})
                 : fn(IntC<0>{}, BoolC<false>{}, BoolC<false>{}, BoolC<true>{});
    if (trk == 1)
      return sec ? fn(IntC<1>{}, BoolC<false>{}, BoolC<true>{}, BoolC<true>{})
                 : fn(IntC<1>{}, BoolC<false>{}, BoolC<false>{}, BoolC<true>{});
    return fn(IntC<2>{}, BoolC<false>{}, BoolC<true>{}, BoolC<true>{});
  }
  if (trk == 2)
    return trace ? fn(IntC<2>{}, BoolC<true>{}, BoolC<true>{}, BoolC<false>{})
                 : fn(IntC<2>{}, BoolC<false>{}, BoolC<true>{}, BoolC<false>{});
  if (trk == 0) {
    if (trace)
      return sec ? fn(IntC<0>{}, BoolC<true>{}, BoolC<true>{}, BoolC<false>{})
                 : fn(IntC<0>{}, BoolC<true>{}, BoolC<false>{}, BoolC<false>{});
    return sec ? fn(IntC<0>{}, BoolC<false>{}, BoolC<true>{}, BoolC<false>{})
               : fn(IntC<0>{}, BoolC<false>{}, BoolC<false>{}, BoolC<false>{});
  }
  if (trace)
    return sec ? fn(IntC<1>{}, BoolC<true>{}, BoolC<true>{}, BoolC<false>{})
               : fn(IntC<1>{}, BoolC<true>{}, BoolC<false>{}, BoolC<false>{});
  return sec ? fn(IntC<1>{}, BoolC<false>{}, BoolC<true>{}, BoolC<false>{})
             : fn(IntC<1>{}, BoolC<false>{}, BoolC<false>{}, BoolC<false>{});
}

// ... x extended geometry (hexagonal lattices, cylinders off the coordinate
// axes): fn(T, TR, SE, SP, EX).
template <class Fn>
static hipError_t dispatch_variant(int trk, bool trace, bool sec, bool split, bool ext, Fn&& fn) {
  if (ext)
    return dispatch_variant_t(trk, trace, sec, split, [&](auto T, auto TR, auto SE, auto SP) {
      return fn(T, TR, SE, SP, BoolC<true>{});
    });
  return dispatch_variant_t(trk, trace, sec, split, [&](auto T, auto TR, auto SE, auto SP) {
    return fn(T, TR, SE, SP, BoolC<false>{});
  });
}

static inline hipError_t event_transport(const DevProblem& P, const KCtx& K, EventBuffers& B,
                                         hipStream_t s, bool sec, bool timing, double* kernel_ms,
                                         int64_t* launches) {
  // the event kernels always carry the extended geometry
  return dispatch_variant_t(P.tracking, K.trace != nullptr, sec, P.precision == ABX_SPLIT,
                          [&](auto T, auto TR, auto SE, auto SP) {
                            return event_transport_t<decltype(T)::value, decltype(TR)::value,
                                                     decltype(SE)::value, decltype(SP)::value>(
                                P, K, B, s, timing, kernel_ms, launches);
                          });
}

// ext: the problem needs the extended-geometry variants (its tables' extended)
static inline hipError_t history_transport(const DevProblem& P, bool ext, const KCtx& K,
                                           hipStream_t s, bool sec, unsigned long long* next,
                                           const abx_tuning& tu) {
  return dispatch_variant(P.tracking, K.trace != nullptr, sec, P.precision == ABX_SPLIT, ext,
                          [&](auto T, auto TR, auto SE, auto SP, auto EX) {
                            return history_transport_t<decltype(T)::value, decltype(TR)::value,
                                                       decltype(SE)::value, decltype(SP)::value,
                                                       decltype(EX)::value>(P, K, s, next, tu);
                          });
}

#endif  // ABX_KERNELS_ONLY
