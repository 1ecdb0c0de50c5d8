// kernels.h - GPU sinks, reductions and the history-based transport kernel.
//
// The history kernel runs one lane per history (the reference's
// `omp for schedule(dynamic)` body, src/delta_tracker.cpp:83-86) and is the
// parity vehicle; the event-based kernels (event.h) share physics.h.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/abx.h"
#include "physics.h"
#include "source.h"

// Per-call device context of a transport launch.
struct KCtx {
  const abx_particle* in;
  int64_t n;
  double kcol;
  int32_t converged;
  int32_t sec_cap;        // secondary-stack entries per slot (0 = none possible)
  SecRec* sec;            // [n][sec_cap]
  // fission staging (k-eigenvalue / modified fixed source)
  abx_banked* stage;
  uint32_t* stage_parent;
  uint32_t* stage_daughter;
  unsigned long long* n_stage;   // staging slots handed out (sites + unused slots)
  unsigned long long* n_hole;    // slots handed out but left unused (marked kHole)
  int64_t stage_cap;
  uint32_t* daughters;    // [n] sites per parent
  double* scores;         // [6] k_col k_abs k_trk k_tot leak mig
  unsigned long long* counters;  // [8] histories collisions virtual flights crossings tally lost secondaries
  int32_t* err;
  unsigned long long* err_hist;
  abx_trace_rec* trace;
  int32_t trace_cap;
  int32_t* n_rec;
  // device-resident generation loop: bank size and k of the previous
  // generation read from device memory (null = use n / kcol)
  const unsigned long long* n_dev;
  const double* kcol_dev;
  // replica mode: the bank holds `nrep` independent k-eigenvalue chains in
  // consecutive history segments rep_start[0..nrep]; a history uses its
  // chain's k (rep_kcol) and scores into its chain's sums (rep_scores,
  // [nrep][8]).  Null outside replica mode.
  const unsigned long long* rep_start;
  const double* rep_kcol;
  double* rep_scores;
  int32_t nrep;
  // extended kernels: each staged fission site's parent record, beside
  // `stage` (null: not kept)
  abx_parent_info* stage_pinfo;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sums six scores and eight counters over the wave, lane 0 adds them in.
__device__ __forceinline__ void flush_scores(const KCtx& K, const Scores6& sc,
                                             const uint32_t c[8]) {
  const double s[6] = {sc.k_col, sc.k_abs, sc.k_trk, sc.k_tot, sc.leak, sc.mig};
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const double t = wave_sum(s[i]);
    if (lane == 0 && t != 0.0) unsafeAtomicAdd(K.scores + i, t);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const unsigned long long t = wave_sum_u64((unsigned long long)c[i]);
    if (lane == 0 && t) atomicAdd(K.counters + i, t);
  }
}

// Replica mode: at most this many chains share a bank; a history's chain
// is the segment of the bank that holds it.
constexpr int kMaxReplicas = 16;

__device__ __forceinline__ int chain_of(const unsigned long long* seg, int nrep,
                                        unsigned long long i) {
  int r = 0;
  while (r + 1 < nrep && i >= seg[r + 1]) r++;
  return r;
}

// Replica mode: a lane's scores go to its block's per-chain sums in LDS
// (`acc`, [kMaxReplicas][6]) when the lane moves on to another chain and at
// the end.  Every lane on the GPU moves on to the next chain at about the
// same time (histories are handed out in order), so per-lane device
// atomics on the six sums of one chain would all serialise at the memory
// side; LDS atomics stay inside the CU.
__device__ __forceinline__ void rep_scores_to_lds(double* acc, int rep, const Scores6& sc) {
  const double s[6] = {sc.k_col, sc.k_abs, sc.k_trk, sc.k_tot, sc.leak, sc.mig};
#pragma unroll
  for (int i = 0; i < 6; i++)
    if (s[i] != 0.0) atomicAdd(acc + 6 * rep + i, s[i]);
}

__device__ __forceinline__ void raise_error(const KCtx& K, int code, uint64_t h) {
  if (atomicCAS(K.err, 0, code) == 0) *K.err_hist = (unsigned long long)h;
}

// stage_parent value of a staging slot that holds no site.
constexpr uint32_t kHole = 0xFFFFFFFFu;

// 1: mesh-tally atomics are aggregated over the lanes of a wavefront that
// score the same bin (GpuSink::tally_add); 0: one atomic per lane.  Off by
// default: on the lattice a wave's lanes almost never score one bin at the
// same time (vector stores per wave 1080 with it off, 1079 on), and the
// match costs ~1 % of kernel time.
#ifndef ABX_TALLY_AGG
#define ABX_TALLY_AGG 0
#endif

// Global-memory side of the physics for one lane.  TRACE compiles the
// per-event trace records in; SEC the per-history secondary stack; SPLIT
// selects the split-precision physics.
// EXT: keeps the previous-collision state and writes each fission site's
// parent record (K->stage_pinfo) for exact cancellation.
template <bool TRACE, bool SEC, bool SPLIT = false, bool EXT = false>
struct GpuSink {
  static constexpr bool split = SPLIT;
  static constexpr bool ext = EXT;
  const KCtx* K;
  const DevProblem* P;
  int64_t slot;   // history index in the input bank
  int64_t sslot;  // secondary-stack slot (the history, or the persistent lane)
  int32_t nsec;
  int32_t n_rec;
  uint32_t c[8];  // per-lane event counters (a lane's counts fit 32 bits)
  // This wave's range [srange[0], srange[1]) of staging slots (LDS), or
  // null: one counter atomic per banking call of the wave.
  unsigned long long* srange;

  // Staging slot for a fission site.  With a wave range, the shared counter
  // is touched once per 64 slots: a counter that every wave of the grid
  // increments at nearly every collision serialises at the memory-side
  // atomic unit.  Called by the lanes that bank now (any subset).
  __device__ unsigned long long stage_slot() {
    if (!srange) return atomicAdd(K->n_stage, 1ull);
    const int lane = threadIdx.x & 63;
    const unsigned long long mask = __ballot(1);
    const int leader = __ffsll((long long)mask) - 1;
    const unsigned long long cnt = (unsigned long long)__popcll(mask);
    const unsigned long long rank = (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
    const unsigned long long wb =
        __hip_atomic_load(srange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    const unsigned long long we =
        __hip_atomic_load(srange + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    const unsigned long long avail = we - wb;
    if (cnt <= avail) {
      if (lane == leader)
        __hip_atomic_store(srange, wb + cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      return wb + rank;
    }
    const unsigned long long grab = cnt - avail > 64ull ? cnt - avail : 64ull;
    unsigned long long nb = 0;
    if (lane == leader) nb = atomicAdd(K->n_stage, grab);
    nb = __shfl(nb, leader, 64);
    if (lane == leader) {
      __hip_atomic_store(srange, nb + (cnt - avail), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WAVEFRONT);
      __hip_atomic_store(srange + 1, nb + grab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    return rank < avail ? wb + rank : nb + (rank - avail);
  }
  // At the wave's end (all lanes): mark the unused rest of its range.
  __device__ void close_stage_range() {
    if (!srange) return;
    const int lane = threadIdx.x & 63;
    const unsigned long long wb =
        __hip_atomic_load(srange, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    const unsigned long long we =
        __hip_atomic_load(srange + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    for (unsigned long long j = wb + (unsigned long long)lane; j < we; j += 64ull)
      if ((int64_t)j < K->stage_cap) K->stage_parent[j] = kHole;
    if (lane == 0 && we > wb) atomicAdd(K->n_hole, we - wb);
  }

  __device__ void tally_add(double* a, double v) {
#if ABX_TALLY_AGG
    // Wavefront aggregation: the lanes scoring the same bin as the first
    // active lane add their scores together (in lane order) and that lane
    // issues one atomic for them; every other lane adds its own.
    const uintptr_t ai = (uintptr_t)a;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)ai);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(ai >> 32));
    const bool same = ai == (((uintptr_t)hi << 32) | lo);
    const unsigned long long grp = __ballot(same);
    if (__popcll(grp) > 1) {
      double s = 0.0;
      // grp is uniform and its lanes are all active, so every shuffle
      // reads an active lane
      for (unsigned long long g = grp; g; g &= g - 1ull) s += __shfl(v, __ffsll(g) - 1);
      if (!same) unsafeAtomicAdd(a, v);
      else if ((threadIdx.x & 63) == __ffsll(grp) - 1) unsafeAtomicAdd(a, s);
    } else {
      unsafeAtomicAdd(a, v);
    }
#else
    unsafeAtomicAdd(a, v);
#endif
    c[5]++;
  }
  __device__ bool sec_push(const SecRec& q) {
    if constexpr (!SEC) {
      return false;
    } else {
      if (nsec >= K->sec_cap) return false;
      K->sec[sslot * K->sec_cap + nsec] = q;
      nsec++;
      return true;
    }
  }
  __device__ bool sec_pop(SecRec& q) {
    if constexpr (!SEC) {
      return false;
    } else {
      if (nsec == 0) return false;
      nsec--;
      q = K->sec[sslot * K->sec_cap + nsec];
      return true;
    }
  }
  __device__ void bank_fission(const Part& p, const V3& u, double E, double w) {
    const unsigned long long i = stage_slot();
    if ((int64_t)i >= K->stage_cap) return;  // counted; host reports capacity
    abx_banked b;
    b.r[0] = p.r.x;
    b.r[1] = p.r.y;
    b.r[2] = p.r.z;
    b.u[0] = u.x;
    b.u[1] = u.y;
    b.u[2] = u.z;
    b.E = E;
    b.wgt = w;
    b.wgt2 = 0.;
    b.parent_history_id = p.hid;
    b.parent_daughter_id = p.daughters;
    b.family_id = p.fam;
    K->stage[i] = b;
    K->stage_parent[i] = (uint32_t)slot;
    K->stage_daughter[i] = p.daughters;
    if constexpr (EXT) {
      // BankedParticle's cancellation fields (src/transporter.cpp:440-456)
      if (K->stage_pinfo) {
        abx_parent_info q;
        q.r1[0] = p.pr.x;
        q.r1[1] = p.pr.y;
        q.r1[2] = p.pr.z;
        q.u1[0] = p.pu.x;
        q.u1[1] = p.pu.y;
        q.u1[2] = p.pu.z;
        q.E1 = p.pE;
        q.E3 = p.E;
        q.Esmp = p.Esmp;
        q.prev_virtual = p.pvirt;
        q.pad_ = 0;
        K->stage_pinfo[i] = q;
      }
    }
  }
  template <class TK>
  __device__ void trace(const Part& p, const TK& t, int ev) {
    if constexpr (TRACE) {
      if (n_rec < K->trace_cap) {
        abx_trace_rec r;
        r.r[0] = p.r.x;
        r.r[1] = p.r.y;
        r.r[2] = p.r.z;
        r.u[0] = p.u.x;
        r.u[1] = p.u.y;
        r.u[2] = p.u.z;
        r.wgt = p.w;
        r.history_id = p.hid;
        r.secondary_id = p.sec_id;
        r.event = ev;
        r.cell = t.cell;
        r.material = t.cell >= 0 ? P->cell[t.cell].fill : -1;
        r.group = p.g;
        r.rng_calls = (int32_t)p.calls;
        K->trace[slot * K->trace_cap + n_rec] = r;
      }
      n_rec++;
    }
  }
  __device__ void error(int code, uint64_t h) { raise_error(*K, code, h); }
  __device__ void count_flight() { c[3]++; }
  __device__ void count_crossing() { c[4]++; }
  __device__ void count_collision() { c[1]++; }
  __device__ void count_virtual() { c[2]++; }
  __device__ void count_lost() { c[6]++; }
  __device__ void count_secondary() { c[7]++; }
  __device__ void init(const KCtx* k, const DevProblem* p, int64_t s) {
    K = k;
    P = p;
    slot = s;
    sslot = s;
    nsec = 0;
    n_rec = 0;
    for (int i = 0; i < 8; i++) c[i] = 0;
    srange = nullptr;
  }
};

__device__ __forceinline__ void load_part(const DevProblem& P, const abx_particle& a, Part& p) {
  p.r = V3{a.r[0], a.r[1], a.r[2]};
  p.u = V3{a.u[0], a.u[1], a.u[2]};
  p.rb = p.r;
  p.E = a.E;
  p.g = energy_index(P, a.E);
  p.w = a.wgt;
  p.w2 = a.wgt2;
  p.rng = a.rng_state;
  p.hid = a.history_id;
  p.fam = a.family_id;
  p.calls = 0;
  p.daughters = 0;
  p.sec_id = 0;
  p.alive = 1;
  // a new Particle's previous state (particle.hpp:231-239)
  p.pr = V3{0., 0., 0.};
  p.pu = V3{1., 0., 0.};
  p.pE = 0.;
  p.Esmp = 0.;
  p.pvirt = 0;
  p.refl = 0;
}

// Dynamic LDS: every lane's cursor pads / tiles / frames, lane-interleaved.
extern __shared__ __align__(16) unsigned char abx_lds[];

template <class TK>
__device__ __forceinline__ void bind_lds_cursor(TK& t, const DevProblem& P) {
  trk_bind(t, abx_lds, blockDim.x, threadIdx.x, P.trk_D, P.trk_L);
}

// History-based transport on a persistent grid.  Each lane carries one
// history at a time and, when it ends, takes the next one from its
// wavefront's private range of history ids, which the wavefront refills
// ABX_WAVE_GRAB ids at a time from a global counter; idle lanes are refilled
// together once ABX_REFILL_MIN of them wait.  Lanes stay busy until the
// bank is drained instead of idling behind the longest history of their
// wavefront.  Per-history results do not depend on which lane runs a
// history: the RNG stream is the history's own and fission sites are
// ordered afterwards by (parent, daughter).
// 32: on the lattice 2.9 % faster than 64 (a wave holds fewer unstarted ids
// when the bank runs dry, so the generation's tail is shorter), 128 is 2 %
// slower; 16 and 24 are no better than 32 and touch the counter more.
#ifndef ABX_WAVE_GRAB
#define ABX_WAVE_GRAB 32ull
#endif
#ifndef ABX_REFILL_MIN
#define ABX_REFILL_MIN 8
#endif

// LDS bytes of the cursors of a block (the staged tables follow them).
__host__ __device__ inline size_t cursor_lds_bytes(int bs, int D, int L) {
  return ((size_t)bs * trk_bytes(D, L) + 15) & ~(size_t)15;
}

// Copies the problem's table blob into LDS after the cursors (whole block)
// and returns the problem reading its tables from there.
__device__ __forceinline__ DevProblem stage_tables(const DevProblem& P) {
  unsigned char* dst = abx_lds + cursor_lds_bytes(blockDim.x, P.trk_D, P.trk_L);
  const int64_t n16 = P.blob_bytes >> 4;
  const uint4* src = (const uint4*)P.blob;
  uint4* d4 = (uint4*)dst;
  for (int64_t i = threadIdx.x; i < n16; i += blockDim.x) d4[i] = src[i];
  __syncthreads();
  return rebase_problem(P, dst);
}

// STAGE: the tables are copied into LDS and read with LDS instructions;
// otherwise read in place (global loads through L2).  A compile-time choice,
// so that each variant's loads carry their address space.
// REP: replica mode (K.rep_start and friends set), a variant of its own so
// that the plain kernels carry none of its code.
// EXT: the problem has hexagonal lattices or cylinders off the coordinate
// axes; likewise its own variant, so that the kernels of other problems
// carry none of that geometry.
template <int TRK, bool TRACE, bool SEC, bool SPLIT, bool STAGE, bool REP = false,
          bool EXT = false>
__global__ void ABX_TKERNEL k_history(DevProblem P0, KCtx K, unsigned long long* next) {
  const DevProblem P = STAGE ? stage_tables(P0) : P0;
  const int64_t n_hist = K.n_dev ? (int64_t)*K.n_dev : K.n;
  const double kcol = K.kcol_dev ? *K.kcol_dev : K.kcol;
  // replica mode: the chains' segments and k, read per history from LDS
  // (nothing per lane is kept: a lane's chain is that of its slot)
  __shared__ unsigned long long s_seg[kMaxReplicas + 1];
  __shared__ double s_kcol[kMaxReplicas];
  __shared__ double s_racc[6 * kMaxReplicas];  // the block's score sums per chain
  if constexpr (REP) {
    if (threadIdx.x <= (unsigned)K.nrep) s_seg[threadIdx.x] = K.rep_start[threadIdx.x];
    if (threadIdx.x < (unsigned)K.nrep) s_kcol[threadIdx.x] = K.rep_kcol[threadIdx.x];
    if (threadIdx.x < 6u * kMaxReplicas) s_racc[threadIdx.x] = 0.0;
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  Scores6 sc{0, 0, 0, 0, 0, 0};
  GpuSink<TRACE, SEC, SPLIT, EXT> s;
  s.init(&K, &P, 0);
  s.sslot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // per-wave fission staging ranges (blocks of at most 1024 lanes)
  __shared__ unsigned long long s_srange[32];
  if (K.stage && K.n_hole) {
    s.srange = s_srange + 2 * (threadIdx.x >> 6);
    if (lane == 0) {
      __hip_atomic_store(s.srange, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      __hip_atomic_store(s.srange + 1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  }
  TrkT<EXT> t;
  bind_lds_cursor(t, P);
  Part p;
  p.alive = 0;
  int m = -1;
  int64_t steps = 0;
  bool done = false;
  const bool converged = K.converged != 0;
  // The wave's private range [wb, we) of history ids: the global counter is
  // touched once per ABX_WAVE_GRAB histories, not at every refill (a
  // counter shared by every wave of the grid is a serialising hot spot).
  unsigned long long wb = 0, we = 0;
  for (;;) {
    const bool need = !p.alive && !done;
    const unsigned long long mask = __ballot(need);
    // Refill once enough lanes are idle (or none has work): starting a
    // history is a long divergent stretch (source particle, descent from
    // the root), paid once per refill whatever the number of lanes.
    if (mask && (__popcll(mask) >= ABX_REFILL_MIN || __ballot(p.alive) == 0ull)) {
      const unsigned long long cnt = (unsigned long long)__popcll(mask);
      const unsigned long long rank = (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
      const unsigned long long avail = we - wb;
      unsigned long long i64;
      if (cnt <= avail) {
        i64 = wb + rank;
        wb += cnt;
      } else {
        const unsigned long long grab = cnt - avail > ABX_WAVE_GRAB ? cnt - avail : ABX_WAVE_GRAB;
        const int leader = __ffsll((long long)mask) - 1;
        unsigned long long nb = 0;
        if (lane == leader) nb = atomicAdd(next, grab);
        nb = __shfl(nb, leader, 64);
        i64 = rank < avail ? wb + rank : nb + (rank - avail);
        wb = nb + (cnt - avail);
        we = nb + grab;
      }
      if (need) {
        const int64_t i = (int64_t)i64;
        if (i < n_hist) {
          if constexpr (REP) {
            // the lane's scores so far belong to its last history's chain
            // (slot 0 and no scores before its first history)
            const int r_old = chain_of(s_seg, K.nrep, (unsigned long long)s.slot);
            if (chain_of(s_seg, K.nrep, (unsigned long long)i) != r_old) {
              rep_scores_to_lds(s_racc, r_old, sc);
              sc = Scores6{0, 0, 0, 0, 0, 0};
            }
          }
          s.slot = i;
          s.nsec = 0;
          s.n_rec = 0;
          steps = 0;
          load_part(P, K.in[i], p);
          s.c[0]++;
          if (!start_history(P, s, p, t, m)) {
            if (K.daughters) K.daughters[i] = 0;
            if (TRACE && K.n_rec) K.n_rec[i] = s.n_rec;
          }
        } else {
          done = true;
        }
      }
    }
    if (__ballot(!done) == 0ull) break;
    if (p.alive) {
      int ev = 0;
      bool hc = step_flight<TRK>(P, s, p, t, m, converged, sc, ev);
      if (p.alive && hc) {
        if constexpr (REP)
          step_collision(P, s, p, m, s_kcol[chain_of(s_seg, K.nrep, (unsigned long long)s.slot)],
                         converged, sc, ev);
        else
          step_collision(P, s, p, m, kcol, converged, sc, ev);
      }
      bool live = step_finish<TRK>(P, s, p, t, m, ev);
      if (live && ++steps > 100000000) {
        s.error(-3, p.hid);
        p.alive = 0;
        live = false;
      }
      if (!live) {
        if (K.daughters) K.daughters[s.slot] = p.daughters;
        if (TRACE && K.n_rec) K.n_rec[s.slot] = s.n_rec;
      }
    }
  }
  s.close_stage_range();
  if constexpr (REP) {
    rep_scores_to_lds(s_racc, chain_of(s_seg, K.nrep, (unsigned long long)s.slot), sc);
    sc = Scores6{0, 0, 0, 0, 0, 0};
    __syncthreads();  // every wave of the block has added its sums
    if (threadIdx.x < 6u * (unsigned)K.nrep && s_racc[threadIdx.x] != 0.0)
      unsafeAtomicAdd(K.rep_scores + 8 * (threadIdx.x / 6) + threadIdx.x % 6,
                      s_racc[threadIdx.x]);
  }
  flush_scores(K, sc, s.c);
}

// Work item of the block-affine batch kernel: histories [first, last) of
// problem `item` (local history ids, read from ctxs[item].in).
struct BatchChunk {
  int32_t item;
  int32_t pad_;
  int64_t first, last;
};

// Many problems in one persistent launch, one problem per block at a time:
// a block takes a chunk of one problem's histories from a global counter,
// stages that problem's tables in LDS when it changes (STAGE), and its lanes
// take the chunk's histories from an LDS counter as in k_history.  The
// problem's tables and context are block-uniform, so they sit in scalar
// registers and the table reads are LDS reads.  A block waits at a chunk's
// end for its longest history; chunks of several thousand histories keep
// that wait small.
template <int TRK, bool SEC, bool SPLIT, bool STAGE, bool EXT = false>
__global__ void ABX_TKERNEL k_batch_chunks(const DevProblem* __restrict__ probs,
                                           const KCtx* __restrict__ ctxs,
                                           const BatchChunk* __restrict__ chunks, int n_chunks,
                                           int D, int L, unsigned* next_chunk) {
  __shared__ int s_chunk;
  __shared__ unsigned long long s_next;
  const int lane = threadIdx.x & 63;
  TrkT<EXT> t;
  trk_bind(t, abx_lds, blockDim.x, threadIdx.x, D, L);
  unsigned char* tables = abx_lds + cursor_lds_bytes(blockDim.x, D, L);
  int staged = -1;
  GpuSink<false, SEC, SPLIT, EXT> s;
  s.init(nullptr, nullptr, 0);
  s.sslot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (;;) {
    if (threadIdx.x == 0) s_chunk = (int)atomicAdd(next_chunk, 1u);
    __syncthreads();
    const int c = s_chunk;
    if (c >= n_chunks) break;
    const BatchChunk ch = chunks[c];
    const int j = ch.item;
    if (threadIdx.x == 0) s_next = (unsigned long long)ch.first;
    if (STAGE && j != staged) {
      const int64_t n16 = probs[j].blob_bytes >> 4;
      const uint4* src = (const uint4*)probs[j].blob;
      uint4* d4 = (uint4*)tables;
      for (int64_t k = threadIdx.x; k < n16; k += blockDim.x) d4[k] = src[k];
      staged = j;
    }
    __syncthreads();
    const DevProblem P = STAGE ? rebase_problem(probs[j], tables) : probs[j];
    const KCtx& K = ctxs[j];
    s.K = &K;
    s.P = &P;
    Scores6 sc{0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; k++) s.c[k] = 0;
    const bool converged = K.converged != 0;
    Part p;
    p.alive = 0;
    int m = -1;
    int64_t steps = 0;
    bool done = false;
    for (;;) {
      const bool need = !p.alive && !done;
      const unsigned long long mask = __ballot(need);
      if (mask && (__popcll(mask) >= ABX_REFILL_MIN || __ballot(p.alive) == 0ull)) {
        const int leader = __ffsll((long long)mask) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(&s_next, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (need) {
          const int64_t i =
              (int64_t)(base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull)));
          if (i < ch.last) {
            s.slot = i;
            s.nsec = 0;
            steps = 0;
            load_part(P, K.in[i], p);
            s.c[0]++;
            if (!start_history(P, s, p, t, m)) {
              if (K.daughters) K.daughters[i] = 0;
            }
          } else {
            done = true;
          }
        }
      }
      if (__ballot(!done) == 0ull) break;
      if (p.alive) {
        int ev = 0;
        bool hc = step_flight<TRK>(P, s, p, t, m, converged, sc, ev);
        if (p.alive && hc) step_collision(P, s, p, m, K.kcol, converged, sc, ev);
        bool live = step_finish<TRK>(P, s, p, t, m, ev);
        if (live && ++steps > 100000000) {
          s.error(-3, p.hid);
          p.alive = 0;
          live = false;
        }
        if (!live && K.daughters) K.daughters[s.slot] = p.daughters;
      }
    }
    flush_scores(K, sc, s.c);
    __syncthreads();
  }
}

// The bookkeeping kernels below are compiled with the host code (abx.hip)
// only, not in the parts of instances.hip.
#ifndef ABX_KERNELS_ONLY

__global__ void k_sample_sources(DevProblem P, DevSources S, uint64_t first, int64_t n,
                                 abx_particle* out, int32_t* err,
                                 unsigned long long* err_hist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TrkT<true> t;
  bind_lds_cursor(t, P);
  abx_particle a;
  const int rc = sample_source(P, S, first + (uint64_t)i, t, a);
  if (rc) {
    if (atomicCAS(err, 0, rc) == 0) *err_hist = first + (uint64_t)i;
    return;
  }
  out[i] = a;
}

// A batch step's source sampling in one launch: job j samples n histories
// from id `first` into `out` with problem probs[prob] and sources srcs[j]
// (SourceDistribution sampling, src/source.cpp:44-90, as k_sample_sources).
// Cursor slices are laid out for the launch's largest capacities (D, L).
struct SampleJob {
  int64_t start;  // prefix over the jobs' histories
  int64_t n;
  uint64_t first;
  abx_particle* out;
  int32_t* err;
  unsigned long long* err_hist;
  int32_t prob;
  int32_t pad_;
};

__global__ void k_sample_jobs(const DevProblem* probs, const DevSources* srcs,
                              const SampleJob* jobs, int njobs, int64_t total, int D, int L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int lo = 0, hi = njobs - 1;  // the last job starting at or before i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].start <= i) lo = mid;
    else hi = mid - 1;
  }
  const SampleJob& J = jobs[lo];
  const int64_t k = i - J.start;
  TrkT<true> t;
  trk_bind(t, abx_lds, blockDim.x, threadIdx.x, D, L);
  abx_particle a;
  const int rc = sample_source(probs[J.prob], srcs[lo], J.first + (uint64_t)k, t, a);
  if (rc) {
    if (atomicCAS(J.err, 0, rc) == 0) *J.err_hist = J.first + (uint64_t)k;
    return;
  }
  J.out[k] = a;
}

// A batch step's tally records in one launch (MeshTally::record_generation
// then clear_generation, src/mesh_tally.cpp:121-152, as k_record_generation):
// job j folds kb generation arrays of one problem into its mean and variance
// in order, the k-th as record number g0 + k + 1, and zeroes them.
struct RecordJob {
  int64_t start;  // prefix over the jobs' bins
  int64_t nb;
  double* gen0;   // first generation array
  double* more;   // the others, nb apart
  double* avg;
  double* var;
  double g0;
  int32_t kb;
  int32_t pad_;
};

__global__ void k_record_jobs(const RecordJob* jobs, int njobs, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].start <= i) lo = mid;
    else hi = mid - 1;
  }
  const RecordJob& J = jobs[lo];
  const int64_t b = i - J.start;
  double a = J.avg[b], v = J.var[b];
  for (int k = 0; k < J.kb; k++) {
    double* gen = k == 0 ? J.gen0 : J.more + (int64_t)(k - 1) * J.nb;
    const double dg = J.g0 + (double)(k + 1);
    const double old_avg = a;
    const double val = gen[b] * 1.0;
    a = old_avg + (val - old_avg) / dg;
    v = v + (((val - old_avg) * (val - a) - (v)) / dg);
    gen[b] = 0.0;
  }
  J.avg[b] = a;
  J.var[b] = v;
}

// ---------------------------------------------------------------- scans
// Exclusive scan of uint32 counts into uint64 offsets, three phases.
__global__ void k_scan_blocks(const uint32_t* in, uint64_t* out, uint64_t* block_sums,
                              int64_t n) {
  __shared__ uint64_t wsum[16];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t v = i < n ? in[i] : 0;
  // inclusive wave scan
  uint64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t acc = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) {
      const uint64_t t = wsum[w];
      wsum[w] = acc;
      acc += t;
    }
    block_sums[blockIdx.x] = acc;
  }
  __syncthreads();
  if (i < n) out[i] = x - v + wsum[wid];
}

__global__ void k_add_block_offsets(uint64_t* out, const uint64_t* block_off, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += block_off[blockIdx.x];
}

// Scatters staged fission sites to parent order.  A slot whose parent is
// not one of the n parents, or whose destination lies past the `total`
// sites, would be a staging slot handed out but never written: it raises
// ABX_E_INVALID in *err instead of writing.
__global__ void k_order_bank(const abx_banked* stage, const uint32_t* parent,
                             const uint32_t* daughter, const uint64_t* offsets, int64_t m,
                             int64_t n, uint64_t total, abx_banked* out, int32_t* err,
                             const abx_parent_info* pin, abx_parent_info* pout) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t q = parent[k];
  if (q == kHole) return;
  if ((int64_t)q >= n) {
    atomicCAS(err, 0, ABX_E_INVALID);
    return;
  }
  const uint64_t dst = offsets[q] + daughter[k];
  if (dst >= total) {
    atomicCAS(err, 0, ABX_E_INVALID);
    return;
  }
  out[dst] = stage[k];
  if (pout) pout[dst] = pin[k];  // the site's parent record (exact cancellation)
}

// MeshTally::record_generation (src/mesh_tally.cpp:121-150), then zero the
// generation array (clear_generation).
__global__ void k_record_generation(double* gen, double* avg, double* var, int64_t n,
                                    double dg, int clear, double multiplier) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double old_avg = avg[i];
  const double val = gen[i] * multiplier;
  const double a = old_avg + (val - old_avg) / dg;
  avg[i] = a;
  double v = var[i];
  v = v + (((val - old_avg) * (val - a) - (v)) / dg);
  var[i] = v;
  if (clear) gen[i] = 0.0;
}

// Zeroes n counters: a kernel, not hipMemsetAsync, so that it can sit in a
// captured graph.
__global__ void k_zero_u64(unsigned long long* a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

__global__ void k_fill(double* a, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// PowerIterator::normalize_weights partial sums (positive / negative weight)
// Grid-stride over the bank, summed per block (waves, then LDS) and added
// once per block: per-wave atomics on the two sums serialise at the memory.
__global__ void k_weight_sums(const abx_banked* b, int64_t n, double* sums) {
  __shared__ double part[2][16];
  double wp = 0., wn = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double w = b[i].wgt;
    if (w > 0.)
      wp += w;
    else
      wn -= w;
  }
  wp = wave_sum(wp);
  wn = wave_sum(wn);
  const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][wv] = wp;
    part[1][wv] = wn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0., c = 0.;
    for (int k = 0; k < nw; k++) {
      a += part[0][k];
      c += part[1][k];
    }
    if (a != 0.) unsafeAtomicAdd(sums, a);
    if (c != 0.) unsafeAtomicAdd(sums + 1, c);
  }
}

// Re-bank (src/power_iterator.cpp:394-404, src/modified_fixed_source.cpp:88-93):
// Particle(r, u, E, w * w_per_part, history_counter++) with a fresh stream.
__global__ void k_rebank(const abx_banked* b, int64_t n, double w_per_part, uint64_t first_hist,
                         uint64_t seed, uint64_t stride, int keep_wgt2, int keep_family,
                         abx_particle* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const abx_banked& s = b[i];
  abx_particle p;
  for (int k = 0; k < 3; k++) {
    p.r[k] = s.r[k];
    p.u[k] = s.u[k];
  }
  p.E = s.E;
  p.wgt = s.wgt * w_per_part;
  p.wgt2 = keep_wgt2 ? s.wgt2 : 0.;
  p.history_id = first_hist + (uint64_t)i;
  p.family_id = keep_family ? s.family_id : 0;
  p.rng_state = pcg_history_state(seed, stride, p.history_id);
  out[i] = p;
}

constexpr int kExPer = 8;                 // values per lane and block
constexpr int kExBlock = 64 * kExPer;     // values a wave stages at a time

// Sums a[k][s0 .. s1) for k < K in index order, one addition after
// another (every lane of the wave gets the sums).  Region kernels run one
// wave per block.  The wave stages 512 consecutive values of each array in
// LDS (coalesced, the next block's loads in flight meanwhile) and every lane
// walks them from LDS with broadcast reads; independent sums interleave.
template <int K>
__device__ void ex_seq_sums(const double* const* a, int64_t s0, int64_t s1, double* acc) {
  __shared__ double buf[K][kExBlock];
  const int lane = threadIdx.x & 63;
  double v[K][kExPer];
  auto fetch = [&](int64_t c) {
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int j = 0; j < kExPer; j++) {
        const int64_t t = c + j * 64 + lane;
        v[k][j] = t < s1 ? a[k][t] : 0.;
      }
  };
  if (s0 < s1) fetch(s0);
  for (int64_t c = s0; c < s1; c += kExBlock) {
    __syncthreads();  // the previous block's walk is over
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int j = 0; j < kExPer; j++) buf[k][j * 64 + lane] = v[k][j];
    __syncthreads();
    if (c + kExBlock < s1) fetch(c + kExBlock);
    const int nblk = (int)(s1 - c < kExBlock ? s1 - c : kExBlock);
    int i = 0;
    for (; i + 8 <= nblk; i += 8) {
#pragma unroll
      for (int u = 0; u < 8; u++)
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += buf[k][i + u];
    }
    for (; i < nblk; i++)
#pragma unroll
      for (int k = 0; k < K; k++) acc[k] += buf[k][i];
  }
}

// ---------------------------------------------------------- cancellation
// ApproximateMeshCancelator (src/approximate_mesh_cancelator.cpp:97-190) on a
// device bank: key = (bin << 32) | index, sorted (each bin's sites in bank
// order), the weights gathered into sorted order, then one wave per bin
// finds the bin's signs and, if the bin holds more than one particle and
// both signs of wgt (wgt2), sums the weights in bank order, one addition
// after another as the reference does, and writes the mean back.
struct DevCancel {
  double lx, ly, lz, dx, dy, dz;
  int32_t nx, ny, nz, ne;  // ne = number of energy bins (1 without edges)
  int32_t n_edges;         // 0 = no energy binning
  const double* edges;
};

__global__ void k_cancel_keys(const abx_banked* b, int64_t n, DevCancel C, uint64_t* keys) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const abx_banked& p = b[t];
  const int i = (int)floor((p.r[0] - C.lx) / C.dx);
  const int j = (int)floor((p.r[1] - C.ly) / C.dy);
  const int k = (int)floor((p.r[2] - C.lz) / C.dz);
  int l = -1;
  if (C.n_edges > 0) {
    for (int e = 0; e < C.n_edges - 1; e++)
      if (C.edges[e] <= p.E && p.E <= C.edges[e + 1]) {
        l = e;
        break;
      }
  } else {
    l = 0;
  }
  uint64_t key = 0xffffffffull;  // outside: its own bin past the last one
  if (!(i < 0 || j < 0 || k < 0 || l < 0) && !(i >= C.nx || j >= C.ny || k >= C.nz || l >= C.ne))
    key = (uint64_t)(l + C.ne * (k + C.nz * (j + C.ny * i)));
  keys[t] = (key << 32) | (uint64_t)t;
}

// Sorted position t: the site's weights in sorted order.
__global__ void k_cancel_gather(const abx_banked* __restrict__ b, const uint64_t* __restrict__ keys,
                                int64_t n, double* __restrict__ w, double* __restrict__ w2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t key = keys[t];
  if ((key >> 32) == 0xffffffffull) return;  // outside the mesh: never read
  const abx_banked& q = b[key & 0xffffffffull];
  w[t] = q.wgt;
  w2[t] = q.wgt2;
}

constexpr int kCancelChunk = 512;  // sorted positions whose bin heads a block takes

// One wave per block.  A block takes the bins that start in its chunks of
// sorted positions (grid-stride), finds each bin's end, its signs, and, for
// a mixed bin of more than one site, its means, then writes them to the
// bin's sites.  Every bin is read and written by one block only, and the
// bank is not read here, so blocks never race.
__global__ void __launch_bounds__(64) k_cancel_bins(const uint64_t* __restrict__ keys, int64_t n,
                                                   const double* __restrict__ w,
                                                   const double* __restrict__ w2,
                                                   abx_banked* __restrict__ b) {
  const int lane = threadIdx.x;
  for (int64_t c0 = (int64_t)blockIdx.x * kCancelChunk; c0 < n;
       c0 += (int64_t)gridDim.x * kCancelChunk) {
    const int64_t c1 = c0 + kCancelChunk < n ? c0 + kCancelChunk : n;
    for (int64_t base = c0; base < c1; base += 64) {
      const int64_t t = base + lane;
      bool head = false;
      if (t < c1) {
        const uint64_t bin = keys[t] >> 32;
        head = bin != 0xffffffffull && (t == 0 || (keys[t - 1] >> 32) != bin);
      }
      uint64_t heads = __ballot(head);
      while (heads) {
        const int64_t h = base + (__ffsll((unsigned long long)heads) - 1);
        heads &= heads - 1;
        const uint64_t bin = keys[h] >> 32;
        int64_t e = h + 1;  // the bin's end: the first later position of another bin
        for (;;) {
          const int64_t u = e + lane;
          const uint64_t out = __ballot(!(u < n && (keys[u] >> 32) == bin));
          if (out) {
            e += __ffsll((unsigned long long)out) - 1;
            break;
          }
          e += 64;
        }
        if (e - h <= 1) continue;
        uint32_t s = 0u;
        for (int64_t u = h + lane; u < e; u += 64) {
          const double x = w[u], y = w2[u];
          if (x > 0.)
            s |= 1u;
          else if (x < 0.)
            s |= 2u;
          if (y > 0.)
            s |= 4u;
          else if (y < 0.)
            s |= 8u;
        }
        const bool m1 = __ballot(s & 1u) != 0 && __ballot(s & 2u) != 0;
        const bool m2 = __ballot(s & 4u) != 0 && __ballot(s & 8u) != 0;
        if (!m1 && !m2) continue;
        double acc[2] = {0., 0.};
        const double* a[2] = {w, w2};
        ex_seq_sums<2>(a, h, e, acc);
        const double N = (double)(e - h);
        const double a1 = acc[0] / N, a2 = acc[1] / N;
        for (int64_t u = h + lane; u < e; u += 64) {
          abx_banked& q = b[keys[u] & 0xffffffffull];
          if (m1) q.wgt = a1;
          if (m2) q.wgt2 = a2;
        }
      }
    }
  }
}

// ------------------------------------------- device-resident generations
// The k-eigenvalue generation loop (PowerIterator::run,
// src/power_iterator.cpp:305-429) with every count and scalar kept on the
// device, so that a generation is a fixed sequence of launches with no host
// round trip and can be captured once as a graph and replayed.  Grids are
// sized for the bank capacity; kernels read the live counts from GenDev.
struct GenDev {
  unsigned long long n;           // histories of the current generation
  unsigned long long n_sites;     // fission sites it banked (after the scan)
  unsigned long long n_stage;     // staging slots handed out by the transport kernel
  unsigned long long n_hole;      // of which left unused
  unsigned long long first_hist;  // first history id of the re-banked generation
  unsigned long long run_hist;    // Simulation::histories_counter
  unsigned long long next;        // persistent-kernel history counter
  unsigned long long err_hist;
  unsigned long long cap;         // bank / staging capacity
  unsigned long long counters[8]; // this generation's event counters
  unsigned long long diag[8];     // summed over the run
  double scores[8];               // this generation's k_col k_abs k_trk k_tot leak mig
  double wsum[2];                 // positive / negative banked weight
  double kcol;                    // Tallies::kcol() for the next transport
  double wpp;                     // nparticles / total weight of the re-bank
  // running statistics of k_col k_abs k_trk k_tot leakage mig (score order)
  double s_avg[6], s_var[6];
  int32_t hist_cap;               // generations the per-generation arrays hold
  int32_t err;
  int32_t gen;                    // recorded (active) generations
  int32_t hist_idx;               // generations run (index into the per-generation arrays)
};

__global__ void k_gen_begin(GenDev* g) {
  const int t = threadIdx.x;
  if (t < 8) {
    g->scores[t] = 0.0;
    g->counters[t] = 0;
  }
  if (t < 2) g->wsum[t] = 0.0;
  if (t == 0) {
    g->n_stage = 0;
    g->n_hole = 0;
    g->next = 0;
  }
}

// Exclusive scan of daughters[0 .. *n) into offsets, first phase (1024 per
// block; blocks past the live count write zero block sums).
__global__ void k_scan_blocks_dev(const uint32_t* in, uint64_t* out, uint64_t* block_sums,
                                  const unsigned long long* n_dev) {
  __shared__ uint64_t wsum[16];
  const int64_t n = (int64_t)*n_dev;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t v = i < n ? in[i] : 0;
  uint64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t acc = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) {
      const uint64_t t = wsum[w];
      wsum[w] = acc;
      acc += t;
    }
    block_sums[blockIdx.x] = acc;
  }
  __syncthreads();
  if (i < n) out[i] = x - v + wsum[wid];
}

// Second phase: one block scans the block sums in place (exclusive) and
// writes the total to *total.
__global__ void k_scan_sums(uint64_t* block_sums, int64_t nb, unsigned long long* total) {
  __shared__ uint64_t part[1024];
  uint64_t carry = 0;
  for (int64_t base = 0; base < nb; base += blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const uint64_t v = i < nb ? block_sums[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
      const uint64_t y = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < nb) block_sums[i] = carry + part[threadIdx.x] - v;
    carry += part[blockDim.x - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void k_add_block_offsets_dev(uint64_t* out, const uint64_t* block_off,
                                        const unsigned long long* n_dev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (int64_t)*n_dev) out[i] += block_off[blockIdx.x];
}

// Scatters the staged sites into parent order and sums their positive and
// negative weights (normalize_weights, src/power_iterator.cpp:538-586):
// grid-stride over the staging slots, one pair of atomics per block.
__global__ void k_order_bank_dev(const abx_banked* stage, const uint32_t* parent,
                                 const uint32_t* daughter, const uint64_t* offsets,
                                 GenDev* g, abx_banked* out) {
  __shared__ double part[2][16];
  const int64_t m = (int64_t)(g->n_stage < g->cap ? g->n_stage : g->cap);
  double wp = 0., wn = 0.;
  if (!g->err) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * blockDim.x) {
      const uint32_t q = parent[k];
      if (q == kHole) continue;
      if ((unsigned long long)q >= g->n) {  // a slot handed out but never written
        atomicCAS(&g->err, 0, ABX_E_INVALID);
        continue;
      }
      const uint64_t dst = offsets[q] + daughter[k];
      if (dst >= g->cap) continue;  // k_gen_end reports the overflow
      const abx_banked b = stage[k];
      out[dst] = b;
      if (b.wgt > 0.)
        wp += b.wgt;
      else
        wn -= b.wgt;
    }
  }
  wp = wave_sum(wp);
  wn = wave_sum(wn);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    part[0][wid] = wp;
    part[1][wid] = wn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0., c = 0.;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      a += part[0][w];
      c += part[1][w];
    }
    if (a != 0.) unsafeAtomicAdd(g->wsum, a);
    if (c != 0.) unsafeAtomicAdd(g->wsum + 1, c);
  }
}

__device__ __forceinline__ void gen_upd(int gen, double x, double& avg, double& var) {
  const double dgen = (double)gen;
  const double avg_old = avg, var_old = var;
  avg = avg_old + (x - avg_old) / dgen;
  if (gen > 1) var = var_old + ((x - avg_old) * (x - avg_old) / dgen) - (var_old / (dgen - 1.));
}

// End of a generation (one thread): the generation's values and the running
// statistics of the six scores (Tallies::calc_gen_values / record_generation,
// src/tallies.cpp:159-203), the re-bank's weight factor (normalize_weights,
// src/power_iterator.cpp:538-586), history ids, the next bank size.  `hist`
// ([6][hist_cap], optional) receives every generation's six values, the
// reference's k_col_vec ... mig_vec.  A device error or an empty bank stops
// the run: the next generations transport nothing.
__global__ void k_gen_end(GenDev* g, double W, int converged, double* hist) {
  if (threadIdx.x != 0) return;
  if (!g->err && g->n_sites != g->n_stage - g->n_hole) g->err = -1;
  if (!g->err && g->n_stage > g->cap) g->err = -5;
  if (!g->err && g->n_sites > g->cap) g->err = -5;
  if (!g->err && g->n_sites == 0) g->err = -8;
  for (int k = 0; k < 8; k++) g->diag[k] += g->counters[k];
  double x[6];
  for (int q = 0; q < 6; q++) x[q] = g->scores[q] / W;
  if (hist && g->hist_idx < g->hist_cap)
    for (int q = 0; q < 6; q++) hist[(int64_t)q * g->hist_cap + g->hist_idx] = x[q];
  g->hist_idx++;
  if (converged && !g->err) {
    g->gen++;
    for (int q = 0; q < 6; q++) gen_upd(g->gen, x[q], g->s_avg[q], g->s_var[q]);
  }
  g->wpp = W / (g->wsum[0] - g->wsum[1]);
  g->first_hist = g->run_hist;
  g->run_hist += g->n_sites;
  g->kcol = x[0];
  g->n = g->err ? 0 : g->n_sites;
}

__device__ __forceinline__ void rebank_one(const abx_banked& s, int64_t i, const GenDev* g,
                                           uint64_t seed, uint64_t stride, abx_particle* out) {
  abx_particle p;
  for (int k = 0; k < 3; k++) {
    p.r[k] = s.r[k];
    p.u[k] = s.u[k];
  }
  p.E = s.E;
  p.wgt = s.wgt * g->wpp;
  p.wgt2 = 0.;
  p.history_id = g->first_hist + (uint64_t)i;
  p.family_id = s.family_id;
  p.rng_state = pcg_history_state(seed, stride, p.history_id);
  out[i] = p;
}

__global__ void k_rebank_dev(const abx_banked* b, const GenDev* g, uint64_t seed, uint64_t stride,
                             abx_particle* out) {
  const int64_t n = (int64_t)g->n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    rebank_one(b[i], i, g, seed, stride, out);
}

// MeshTally::record_generation + clear_generation with the generation
// count read on the device (after k_gen_end incremented it).
__global__ void k_record_generation_dev(double* gen, double* avg, double* var, int64_t n,
                                        const GenDev* g, int32_t g0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dg = (double)(g0 + g->gen);
  const double old_avg = avg[i];
  const double val = gen[i] * 1.0;
  const double a = old_avg + (val - old_avg) / dg;
  avg[i] = a;
  double v = var[i];
  v = v + (((val - old_avg) * (val - a) - (v)) / dg);
  var[i] = v;
  gen[i] = 0.0;
}

// ------------------------------------------- replicas: chains in one bank
// `nrep` independent k-eigenvalue chains of one problem (PowerIterator::run,
// src/power_iterator.cpp:305-429, once per chain with its own seed) share
// each generation's launches: their banks are consecutive segments of one
// bank, the transport kernel runs all their histories, and the bookkeeping
// below keeps every chain's k, weight normalisation, history ids and RNG
// seed apart, so that each chain transports exactly the histories it
// transports alone.  The mesh tallies hold the sum of the chains; a
// generation records their mean.
struct GenRep {
  unsigned long long n;          // histories of the generation, all chains
  unsigned long long n_sites;    // fission sites banked, all chains
  unsigned long long n_stage, n_hole, next, err_hist, cap;
  unsigned long long counters[8], diag[8];
  unsigned long long seg[kMaxReplicas + 1];       // chain segments of the bank
  unsigned long long site_seg[kMaxReplicas + 1];  // chain segments of the sites
  unsigned long long first_hist[kMaxReplicas], run_hist[kMaxReplicas];
  unsigned long long seed[kMaxReplicas];
  double scores[kMaxReplicas][8];  // per chain: k_col k_abs k_trk k_tot leak mig
  double wsum[kMaxReplicas][2];    // per chain: positive / negative banked weight
  double kcol[kMaxReplicas];       // per chain: Tallies::kcol() for the next transport
  double wpp[kMaxReplicas];        // per chain: nparticles / total weight
  double s_avg[6], s_var[6];       // running statistics of the chains' mean
  int32_t nrep, hist_cap, err, gen, hist_idx, pad_;
};

__global__ void k_rep_begin(GenRep* g) {
  const int t = threadIdx.x;
  if (t < 8 * kMaxReplicas) (&g->scores[0][0])[t] = 0.0;
  if (t < 2 * kMaxReplicas) (&g->wsum[0][0])[t] = 0.0;
  if (t < 8) g->counters[t] = 0;
  if (t == 0) {
    g->n_stage = 0;
    g->n_hole = 0;
    g->next = 0;
  }
}

// Chain segments of the ordered sites: the exclusive scan of the per-parent
// counts at each chain's first parent (the total past the last).
__global__ void k_rep_sites(GenRep* g, const uint64_t* offsets) {
  if (threadIdx.x != 0) return;
  for (int r = 0; r <= g->nrep; r++)
    g->site_seg[r] = g->seg[r] < g->n ? offsets[g->seg[r]] : g->n_sites;
}

// k_order_bank_dev with the weight sums kept per chain.
__global__ void k_rep_order(const abx_banked* stage, const uint32_t* parent,
                            const uint32_t* daughter, const uint64_t* offsets, GenRep* g,
                            abx_banked* out) {
  __shared__ double part[kMaxReplicas][2];
  if (threadIdx.x < 2 * kMaxReplicas) (&part[0][0])[threadIdx.x] = 0.0;
  __syncthreads();
  const int64_t m = (int64_t)(g->n_stage < g->cap ? g->n_stage : g->cap);
  if (!g->err) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * blockDim.x) {
      const uint32_t q = parent[k];
      if (q == kHole) continue;
      if ((unsigned long long)q >= g->n) {  // a slot handed out but never written
        atomicCAS(&g->err, 0, ABX_E_INVALID);
        continue;
      }
      const uint64_t dst = offsets[q] + daughter[k];
      if (dst >= g->cap) continue;  // k_rep_end reports the overflow
      const abx_banked b = stage[k];
      out[dst] = b;
      const int r = chain_of(g->seg, g->nrep, q);
      if (b.wgt > 0.)
        atomicAdd(&part[r][0], b.wgt);
      else
        atomicAdd(&part[r][1], -b.wgt);
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * g->nrep) {
    const double v = (&part[0][0])[threadIdx.x];
    if (v != 0.) unsafeAtomicAdd(&g->wsum[0][0] + threadIdx.x, v);
  }
}

// End of a generation for every chain (one thread): each chain's values,
// k, weight factor and history ids; the mean of the chains' six values joins
// the per-generation vectors (`hist`, [6][hist_cap]) and the running
// statistics; `khist` ([nrep][hist_cap]) receives every chain's k_col.
__global__ void k_rep_end(GenRep* g, double W, int converged, double* hist, double* khist) {
  if (threadIdx.x != 0) return;
  const int R = g->nrep;
  if (!g->err && g->n_sites != g->n_stage - g->n_hole) g->err = -1;
  if (!g->err && g->n_stage > g->cap) g->err = -5;
  if (!g->err && g->n_sites > g->cap) g->err = -5;
  for (int r = 0; r < R; r++)
    if (!g->err && g->site_seg[r + 1] == g->site_seg[r]) g->err = -8;
  for (int k = 0; k < 8; k++) g->diag[k] += g->counters[k];
  double x[6] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < R; r++) {
    for (int q = 0; q < 6; q++) x[q] += g->scores[r][q] / W;
    if (khist && g->hist_idx < g->hist_cap)
      khist[(int64_t)r * g->hist_cap + g->hist_idx] = g->scores[r][0] / W;
  }
  for (int q = 0; q < 6; q++) x[q] /= (double)R;
  if (hist && g->hist_idx < g->hist_cap)
    for (int q = 0; q < 6; q++) hist[(int64_t)q * g->hist_cap + g->hist_idx] = x[q];
  g->hist_idx++;
  if (converged && !g->err) {
    g->gen++;
    for (int q = 0; q < 6; q++) gen_upd(g->gen, x[q], g->s_avg[q], g->s_var[q]);
  }
  for (int r = 0; r < R; r++) {
    g->wpp[r] = W / (g->wsum[r][0] - g->wsum[r][1]);
    g->first_hist[r] = g->run_hist[r];
    g->run_hist[r] += g->site_seg[r + 1] - g->site_seg[r];
    g->kcol[r] = g->scores[r][0] / W;
  }
  for (int r = 0; r <= R; r++) g->seg[r] = g->site_seg[r];
  g->n = g->err ? 0 : g->n_sites;
}

// Re-bank of every chain's sites with the chain's weight factor, history
// ids and seed.
__global__ void k_rebank_rep(const abx_banked* b, const GenRep* g, uint64_t stride,
                             abx_particle* out) {
  const int64_t n = (int64_t)g->n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = chain_of(g->seg, g->nrep, (unsigned long long)i);
    const abx_banked& s = b[i];
    abx_particle p;
    for (int k = 0; k < 3; k++) {
      p.r[k] = s.r[k];
      p.u[k] = s.u[k];
    }
    p.E = s.E;
    p.wgt = s.wgt * g->wpp[r];
    p.wgt2 = 0.;
    p.history_id = g->first_hist[r] + (uint64_t)(i - (int64_t)g->seg[r]);
    p.family_id = s.family_id;
    p.rng_state = pcg_history_state(g->seed[r], stride, p.history_id);
    out[i] = p;
  }
}

// A chain started from a shared bank: the bank's particles become the
// chain's histories 0..n-1 on its own stream (`dst` may be `src`).
__global__ void k_fork_chain(const abx_particle* src, int64_t n, uint64_t seed, uint64_t stride,
                             abx_particle* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  abx_particle q = src[i];
  q.history_id = (uint64_t)i;
  q.rng_state = pcg_history_state(seed, stride, q.history_id);
  dst[i] = q;
}

// MeshTally::record_generation of the chains' mean (the generation array
// holds their sum), with the generation count read on the device.
__global__ void k_record_rep(double* gen, double* avg, double* var, int64_t n, const GenRep* g,
                             double mult) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dg = (double)g->gen;
  const double old_avg = avg[i];
  const double val = gen[i] * mult;
  const double a = old_avg + (val - old_avg) / dg;
  avg[i] = a;
  double v = var[i];
  v = v + (((val - old_avg) * (val - a) - (v)) / dg);
  var[i] = v;
  gen[i] = 0.0;
}

#endif  // ABX_KERNELS_ONLY
