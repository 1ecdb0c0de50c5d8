// exact_kernels.h - ExactMGCancelator / BasicExactMGCancelator
// (src/exact_mg_cancelator.cpp, src/basic_exact_mg_cancelator.cpp) on a
// device-resident fission bank and its parent records.
//
// Pipeline (the host drives it, abx.hip); "site" kernels run one lane per
// site in sorted order, "region" kernels one wave per region:
//   k_ex_keys      site: region number (exact.h) << 32 | bank index
//   radix sort     regions in order, each region's sites in bank order
//   k_ex_heads     region starts; an exclusive scan numbers the regions
//   k_ex_bins      one record per region: its range of sorted positions
//   k_ex_site      site: its weights in sorted order, the signs present
//   k_ex_region    region: mixed signs, W and W2 (basic average-f)
//   k_ex_samples   one lane per region: the n_samples points (Sobol or pcg32)
//   k_ex_f         site: <f>, <1/f> and the C terms
//   k_ex_sums      region: sum_c, sum_c_wgt, sum_c_wgt2
//   k_ex_p         site: beta, its own f, the kept and uniform shares of
//                  wgt and wgt2, and where a region's pass stops
//   k_ex_uniform   region: the uniform weights up to the stops
//   k_ex_apply     site: the new weights
//   k_ex_new       one block: the uniform particles, in region order, from
//                  the one global pcg32 stream
// The reference forms its per-region sums in bank order, one addition after
// another; the region kernels do the same additions in the same order (a
// wave loads 512 consecutive values, then adds them one at a time), so the
// weights and new particles are bit-identical to a serial pass.  Everything
// else is independent per site and runs in parallel.
#pragma once

#include "exact.h"
#include "kernels.h"

struct ExBinD {
  int64_t start, end;  // sorted positions [start, end)
  uint64_t reg;        // region number
  uint32_t sign;       // bit 0/1: some wgt > 0 / < 0, bit 2/3: the same for wgt2
  int32_t m1, m2;      // wgt / wgt2 of both signs present (cancel_bin runs)
  int32_t avg;         // <f> and <1/f> are computed for it
  int32_t can_cancel;
  int32_t pad_;
  int64_t stop1, stop2;  // first sorted position whose P_p / P_u ends the pass
  double W, W2;
  double sum_c, sum_c_wgt, sum_c_wgt2;
  double uw, uw2;      // uniform weights
  uint32_t n_new;      // uniform particles
  int32_t pad2_;
};

struct ExSample {
  double x, y, z;
  int32_t g;
  int32_t pad_;
};

// Per-site arrays, in sorted order.
struct ExSiteArrays {
  double *w, *w2;          // weights
  double *fv, *fi;         // <f>, <1/f>
  double *cc, *cw, *cw2;   // C, C * wgt, C * wgt2
  double *kp1, *up1;       // wgt * P_p, wgt * P_u
  double *kp2, *up2;       // the same for wgt2
};

__device__ __forceinline__ uint64_t ex_reg_of(uint64_t key) { return key >> 32; }
__device__ __forceinline__ uint32_t ex_site_of(uint64_t key) { return (uint32_t)key; }

__global__ void k_ex_keys(DevProblem P, DevExact E, const abx_banked* __restrict__ bank,
                          const abx_parent_info* __restrict__ par, int64_t n,
                          uint64_t* __restrict__ keys, int32_t* err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TrkT<true> t;
  bind_lds_cursor(t, P);
  const abx_banked b = bank[i];
  uint64_t reg = kExNone;
  ExKeyD K;
  if (ex_site_key(E, P, b, par[i].prev_virtual, K)) {
    const int mat = ex_material(P, t, V3{b.r[0], b.r[1], b.r[2]});
    if (mat < 0)
      atomicCAS(err, 0, ABX_E_INVALID);  // a fission site in no cell
    else
      reg = ex_region(E, K.i, K.j, K.k, K.e, mat);
  }
  keys[i] = (reg << 32) | (uint64_t)i;
}

__global__ void k_ex_heads(const uint64_t* __restrict__ keys, int64_t n,
                           uint32_t* __restrict__ head) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t r = ex_reg_of(keys[t]);
  head[t] = (r != kExNone && (t == 0 || ex_reg_of(keys[t - 1]) != r)) ? 1u : 0u;
}

// Region records: range and number, everything else cleared.
__global__ void k_ex_bins(const uint64_t* __restrict__ keys, int64_t n,
                          const uint32_t* __restrict__ head, const uint64_t* __restrict__ offs,
                          ExBinD* __restrict__ bins) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t r = ex_reg_of(keys[t]);
  if (r == kExNone) return;
  const uint64_t b = offs[t] + head[t] - 1u;
  if (head[t]) {
    ExBinD& B = bins[b];
    B.start = t;
    B.reg = r;
    B.sign = 0u;
    B.m1 = B.m2 = B.avg = 0;
    B.can_cancel = 1;
    B.W = B.W2 = B.sum_c = B.sum_c_wgt = B.sum_c_wgt2 = B.uw = B.uw2 = 0.;
    B.n_new = 0u;
  }
  if (t == n - 1 || ex_reg_of(keys[t + 1]) != r) bins[b].end = t + 1;
}

__device__ __forceinline__ uint64_t ex_bin_of(const uint32_t* head, const uint64_t* offs,
                                              int64_t t) {
  return offs[t] + head[t] - 1u;
}

// The sites' weights in sorted order and the signs of each region.
__global__ void k_ex_site(const abx_banked* __restrict__ bank, const uint64_t* __restrict__ keys,
                          int64_t n, const uint32_t* __restrict__ head,
                          const uint64_t* __restrict__ offs, ExBinD* __restrict__ bins,
                          ExSiteArrays A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t s = 0u;
  uint64_t bin = ~0ull;  // no region
  if (t < n) {
    const uint64_t key = keys[t];
    if (ex_reg_of(key) != kExNone) {
      const abx_banked& q = bank[ex_site_of(key)];
      const double w = q.wgt, w2 = q.wgt2;
      A.w[t] = w;
      A.w2[t] = w2;
      if (w > 0.)
        s |= 1u;
      else if (w < 0.)
        s |= 2u;
      if (w2 > 0.)
        s |= 4u;
      else if (w2 < 0.)
        s |= 8u;
      bin = ex_bin_of(head, offs, t);
    }
  }
  // sites are sorted by region: OR the wave's runs of one region together
  // (inclusive segmented scan), one atomic per run
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t os = __shfl_up(s, (unsigned)d, 64);
    const uint64_t ob = __shfl_up(bin, (unsigned)d, 64);
    if (lane >= d && ob == bin) s |= os;
  }
  const uint64_t nb_next = __shfl_down(bin, 1u, 64);
  const bool last = lane == 63 || nb_next != bin;
  if (bin != ~0ull && last && s) atomicOr(&bins[bin].sign, s);
}

// add_particle's W, W2 (in bank order) and the cancel_bin / averages
// conditions of perform_cancellation (exact :405-517, basic :471-562).
__global__ void __launch_bounds__(64) k_ex_region(DevExact E, ExBinD* __restrict__ bins,
                                                 int64_t nb, ExSiteArrays A) {
  const int64_t b = blockIdx.x;  // one wave per block and region
  if (b >= nb) return;
  ExBinD& B = bins[b];
  const int64_t s0 = B.start, s1 = B.end;
  const uint32_t sg = B.sign;
  const bool m1 = (sg & 3u) == 3u, m2 = (sg & 12u) == 12u;
  double acc[2] = {0., 0.};
  const bool basic = E.kind == ABX_BASIC_EXACT_MG;
  if (basic && E.beta == ABX_BETA_AVERAGE_F) {  // the only reader of W and W2
    const double* a[2] = {A.w, A.w2};
    ex_seq_sums<2>(a, s0, s1, acc);
  }
  if (threadIdx.x == 0) {
    const bool many = s1 - s0 > 1;
    B.m1 = many && m1;
    B.m2 = many && m2;
    const bool needs = !basic || E.beta == ABX_BETA_AVERAGE_F || E.beta == ABX_BETA_AVERAGE_G;
    B.avg = many && (m1 || m2) && needs;
    B.W = acc[0];
    B.W2 = acc[1];
    B.stop1 = s1;
    B.stop2 = s1;
  }
}

// The points of each region that needs averages (compute_averages' loop,
// exact :303-327; get_averages(_sobol), basic :243-407): Sobol index from 0
// per region, or the region's own pcg32 stream (the global one advanced past
// the draws every earlier region may take, basic :475-489), at most 100
// candidates per point.
__global__ void k_ex_samples(DevProblem P, DevExact E, ExBinD* __restrict__ bins, int64_t nb,
                             uint64_t grng, uint64_t max_rn, ExSample* __restrict__ smp) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  TrkT<true> t;
  bind_lds_cursor(t, P);
  ExBinD& B = bins[b];
  if (!B.avg) return;
  const ExKeyD K = ex_decode(E, B.reg);
  const bool exact = E.kind == ABX_EXACT_MG;
  const bool use_sobol = exact || E.sobol;
  uint64_t rs = use_sobol ? 0u : pcg_advance(grng, (uint64_t)B.start * max_rn);
  uint64_t si = 0;
  for (int j = 0; j < E.n_samples; j++) {
    bool ok = false;
    V3 r{0., 0., 0.};
    for (int tries = 0; tries < 100; tries++) {
      double x0, x1, x2;
      if (use_sobol) {
        x0 = sobol_at(E.sobolM, si, 0);
        x1 = sobol_at(E.sobolM, si, 1);
        x2 = sobol_at(E.sobolM, si, 2);
        si++;
      } else {
        x0 = rng_rand(rs);
        x1 = rng_rand(rs);
        x2 = rng_rand(rs);
      }
      r = ex_point(E, K, x0, x1, x2);
      if (ex_material(P, t, r) == K.mat) {
        ok = true;
        break;
      }
    }
    if (!ok) {
      B.can_cancel = 0;
      return;
    }
    int32_t g = 0;
    if (exact && E.chi_matrix) {  // sample_point (:282-289): 4th coordinate of the last index
      const double xi = sobol_at(E.sobolM, si - 1, 3);
      const int nbins = E.gb_off[K.e + 1] - E.gb_off[K.e];
      const int gi = (int)floor(xi * (double)nbins);
      g = E.gb[E.gb_off[K.e] + gi];
    }
    ExSample& S = smp[b * E.n_samples + j];
    S.x = r.x;
    S.y = r.y;
    S.z = r.z;
    S.g = g;
  }
}

// <f> and <1/f> of each site of a region with averages (compute_averages
// :329-341; get_averages :262-283) and its C terms (:343-364).  The exact
// cancelator gives up on the region at the first f of zero.
__global__ void k_ex_f(DevProblem P, DevExact E, const abx_parent_info* __restrict__ par,
                       const uint64_t* __restrict__ keys, int64_t n,
                       const uint32_t* __restrict__ head, const uint64_t* __restrict__ offs,
                       ExBinD* __restrict__ bins, const ExSample* __restrict__ smp,
                       ExSiteArrays A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t key = keys[t];
  if (ex_reg_of(key) == kExNone) return;
  const uint64_t b = ex_bin_of(head, offs, t);
  ExBinD& B = bins[b];
  if (!B.avg || !__hip_atomic_load(&B.can_cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    return;
  const ExKeyD K = ex_decode(E, B.reg);
  const abx_parent_info q = par[ex_site_of(key)];
  const V3 r1{q.r1[0], q.r1[1], q.r1[2]};
  const ExSample* S = smp + b * E.n_samples;
  double sf = 0., sfi = 0.;
  if (E.kind == ABX_EXACT_MG) {
    const V3 u1{q.u1[0], q.u1[1], q.u1[2]};
    const int g1 = ex_group(P, q.E1), g3 = ex_group(P, q.E3);
    for (int j = 0; j < E.n_samples; j++) {
      const double f =
          ex_exact_f(E, P, K.mat, r1, u1, g1, g3, V3{S[j].x, S[j].y, S[j].z}, S[j].g, q.Esmp);
      if (f == 0.) {
        __hip_atomic_store(&B.can_cancel, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      sf += f;
      sfi += 1. / f;
    }
  } else {
    for (int j = 0; j < E.n_samples; j++) {
      const double f = ex_basic_f(V3{S[j].x, S[j].y, S[j].z}, r1, q.Esmp);
      sf += f;
      sfi += 1. / f;
    }
  }
  const double fv = sf / (double)E.n_samples, fi = sfi / (double)E.n_samples;
  A.fv[t] = fv;
  A.fi[t] = fi;
  const double c = ex_c(fv, fi);
  A.cc[t] = c;
  A.cw[t] = c * A.w[t];
  A.cw2[t] = c * A.w2[t];
}

// sum_c, sum_c_wgt and sum_c_wgt2 of compute_averages (:343-364), in bank
// order, for the regions whose beta uses them.
// sum_c_wgt (sum_c_wgt2) is only read by the wgt (wgt2) pass.
__global__ void __launch_bounds__(64) k_ex_sums(DevExact E, ExBinD* __restrict__ bins, int64_t nb,
                                               ExSiteArrays A) {
  const int64_t b = blockIdx.x;
  if (b >= nb) return;
  ExBinD& B = bins[b];
  const bool uses = E.kind == ABX_EXACT_MG || E.beta == ABX_BETA_AVERAGE_G;
  if (!(B.avg && B.can_cancel && uses)) return;
  double acc[3] = {0., 0., 0.};
  if (B.m1 && B.m2) {
    const double* a[3] = {A.cc, A.cw, A.cw2};
    ex_seq_sums<3>(a, B.start, B.end, acc);
  } else {
    const double* a[2] = {A.cc, B.m1 ? A.cw : A.cw2};
    ex_seq_sums<2>(a, B.start, B.end, acc);
    if (!B.m1) acc[2] = acc[1];
  }
  if (threadIdx.x == 0) {
    B.sum_c = acc[0];
    B.sum_c_wgt = B.m1 ? acc[1] : 0.;
    B.sum_c_wgt2 = B.m2 ? acc[2] : 0.;
  }
}

// cancel_bin for wgt and wgt2 (exact :367-403, basic :438-468), per site:
// beta (get_beta), the site's own f, the kept share wgt * P_p and the
// uniform share wgt * P_u.  A non-finite P_p or P_u, or (basic) a zero
// weight, ends the region's pass there, as the reference's `return` does:
// the earliest such position is the region's stop.
__global__ void k_ex_p(DevProblem P, DevExact E, const abx_banked* __restrict__ bank,
                       const abx_parent_info* __restrict__ par, const uint64_t* __restrict__ keys,
                       int64_t n, const uint32_t* __restrict__ head,
                       const uint64_t* __restrict__ offs, ExBinD* __restrict__ bins,
                       ExSiteArrays A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t key = keys[t];
  if (ex_reg_of(key) == kExNone) return;
  ExBinD& B = bins[ex_bin_of(head, offs, t)];
  if (!B.m1 && !B.m2) return;
  const bool basic = E.kind == ABX_BASIC_EXACT_MG;
  const bool can = B.can_cancel != 0;
  const uint32_t site = ex_site_of(key);
  const abx_banked s = bank[site];
  const abx_parent_info q = par[site];
  const V3 r1{q.r1[0], q.r1[1], q.r1[2]};
  const V3 r4{s.r[0], s.r[1], s.r[2]};
  double f;  // get_f at the site itself
  if (basic) {
    f = ex_basic_f(r4, r1, q.Esmp);
  } else {
    const V3 u1{q.u1[0], q.u1[1], q.u1[2]};
    f = ex_exact_f(E, P, ex_decode(E, B.reg).mat, r1, u1, ex_group(P, q.E1), ex_group(P, q.E3),
                   r4, ex_group(P, s.E), q.Esmp);
  }
  double Bmin = 0.;
  if (basic && can && E.beta == ABX_BETA_MINIMUM) Bmin = ex_basic_min_f(E, ex_decode(E, B.reg), r1, q.Esmp);
  for (int pass = 0; pass < 2; pass++) {
    const bool first = pass == 0;
    if (!(first ? B.m1 : B.m2)) continue;
    const double wgt = first ? s.wgt : s.wgt2;
    bool stop = false;
    double Pp = 0., Pu = 0.;
    if (basic && wgt == 0.) {
      stop = true;
    } else {
      double Bt = 0.;  // get_beta
      if (can) {
        if (!basic || E.beta == ABX_BETA_AVERAGE_G) {
          const double S = (first ? B.sum_c_wgt : B.sum_c_wgt2) / (1. + B.sum_c);
          const double fv = A.fv[t], f_inv = A.fi[t];
          Bt = fv * (1. / (2. * fv * f_inv - 1.)) * (1. - (S / wgt));
        } else if (E.beta == ABX_BETA_MINIMUM) {
          Bt = Bmin;
        } else if (E.beta == ABX_BETA_AVERAGE_F) {
          const double N = (double)(B.end - B.start);
          const double W = first ? B.W : B.W2;
          Bt = A.fv[t] * (1. - ((W) / ((N + 1.) * wgt)));
        }
      }
      Pp = (f - Bt) / f;
      Pu = Bt / f;
      if (isinf(Pu) || isinf(Pp) || isnan(Pu) || isnan(Pp)) stop = true;
    }
    if (stop) {
      atomicMin((unsigned long long*)(first ? &B.stop1 : &B.stop2), (unsigned long long)t);
    } else if (first) {
      A.kp1[t] = wgt * Pp;
      A.up1[t] = wgt * Pu;
    } else {
      A.kp2[t] = wgt * Pp;
      A.up2[t] = wgt * Pu;
    }
  }
}

// The uniform weights (cancel_bin's uniform_wgt sums, in bank order up to
// each pass's stop) and the number of uniform particles
// (get_new_particles: ceil(max(|uniform_wgt|, |uniform_wgt2|))).
__global__ void __launch_bounds__(64) k_ex_uniform(ExBinD* __restrict__ bins, int64_t nb,
                                                  ExSiteArrays A) {
  const int64_t b = blockIdx.x;
  if (b >= nb) return;
  ExBinD& B = bins[b];
  double u1 = 0., u2 = 0.;
  if (B.m1) {
    const double* a[1] = {A.up1};
    ex_seq_sums<1>(a, B.start, B.stop1, &u1);
  }
  if (B.m2) {
    const double* a[1] = {A.up2};
    ex_seq_sums<1>(a, B.start, B.stop2, &u2);
  }
  if (threadIdx.x == 0) {
    B.uw = u1;
    B.uw2 = u2;
    const double x = fabs(u1), y = fabs(u2);
    B.n_new = (uint32_t)ceil(x < y ? y : x);
  }
}

__global__ void k_ex_apply(abx_banked* __restrict__ bank, const uint64_t* __restrict__ keys,
                           int64_t n, const uint32_t* __restrict__ head,
                           const uint64_t* __restrict__ offs, const ExBinD* __restrict__ bins,
                           ExSiteArrays A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t key = keys[t];
  if (ex_reg_of(key) == kExNone) return;
  const ExBinD& B = bins[ex_bin_of(head, offs, t)];
  abx_banked& s = bank[ex_site_of(key)];
  if (B.m1 && t < B.stop1) s.wgt = A.kp1[t];
  if (B.m2 && t < B.stop2) s.wgt2 = A.kp2[t];
}

__global__ void k_ex_counts(const ExBinD* __restrict__ bins, int64_t nb, uint32_t* __restrict__ c) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) c[b] = bins[b].n_new;
}

// get_new_particles (exact :519-605, basic :564-611): each region's uniform
// particles in region order from the one pcg32 stream, each at the first of
// at most 100 candidate points (three numbers each) in the region's
// material, followed by its energy and direction draws (ex_tail_steps).
// One block.  A round fixes the stream's state
// o at the next unresolved particle and tests, side by side, the candidate
// that would start at each of the next kExWindow numbers; one lane then
// walks that table as the reference's loop consumes the stream (a rejected
// candidate moves on three numbers; an accepted one ends the particle,
// whose energy and direction take the next numbers) and the particles it
// resolved are written side by side.  The window holds 100 candidates of
// the round's first particle, so every round resolves at least one.
constexpr int kExWindow = 512;

__global__ void __launch_bounds__(256) k_ex_new(DevProblem P, DevExact E,
                                                const ExBinD* __restrict__ bins, int64_t nb,
                                                const uint64_t* __restrict__ first_new,
                                                uint64_t grng, abx_banked* __restrict__ out,
                                                uint64_t* __restrict__ grng_out, int32_t* err) {
  __shared__ unsigned char s_acc[kExWindow];
  // candidate slot of each resolved particle: a particle takes at least five
  // numbers (three for its position, two or more after it)
  constexpr int kCand = kExWindow / 5 + 1;
  __shared__ int s_cand[kCand];
  __shared__ int s_np, s_next, s_fail;
  TrkT<true> t;
  bind_lds_cursor(t, P);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tail = (int)(ex_tail_steps(E) / 2u);  // numbers after the position
  uint64_t o = grng;  // every lane keeps the same copy
  for (int64_t b = 0; b < nb; b++) {
    const ExBinD B = bins[b];
    const uint32_t N = B.n_new;
    if (N == 0) continue;
    const double w = B.uw / (double)N, w2 = B.uw2 / (double)N;
    const ExKeyD K = ex_decode(E, B.reg);
    abx_banked* dst = out + first_new[b];
    uint32_t j0 = 0;
    while (j0 < N) {
      for (int k = tid; k < kExWindow; k += nt) {
        uint64_t s = pcg_advance(o, 2u * (uint64_t)k);
        const double x0 = rng_rand(s), x1 = rng_rand(s), x2 = rng_rand(s);
        s_acc[k] = ex_material(P, t, ex_point(E, K, x0, x1, x2)) == K.mat ? 1 : 0;
      }
      __syncthreads();
      if (tid == 0) {
        int q = 0, np = 0, fail = 0;
        while (j0 + (uint32_t)np < N && q < kExWindow && np < kCand) {
          int k = q, tries = 0;
          while (tries < 100 && k < kExWindow && !s_acc[k]) {
            k += 3;
            tries++;
          }
          if (tries == 100) {
            fail = 1;
            break;
          }
          if (k >= kExWindow) break;  // its candidates run past the window: next round
          s_cand[np++] = k;
          q = k + 3 + tail;
        }
        s_np = np;
        s_next = q;
        s_fail = fail;
      }
      __syncthreads();
      const int np = s_np, next = s_next;
      if (s_fail && np == 0) {  // "Couldn't sample position for uniform particle."
        if (tid == 0) atomicCAS(err, 0, ABX_E_INVALID);
        return;
      }
      for (int i = tid; i < np; i += nt) {
        uint64_t s = pcg_advance(o, 2u * (uint64_t)s_cand[i]);
        const double x0 = rng_rand(s), x1 = rng_rand(s), x2 = rng_rand(s);
        const V3 r = ex_point(E, K, x0, x1, x2);
        abx_banked u{};
        u.r[0] = r.x;
        u.r[1] = r.y;
        u.r[2] = r.z;
        ex_new_tail(E, P, K, s, u);
        u.wgt = w;
        u.wgt2 = w2;
        dst[j0 + (uint32_t)i] = u;
      }
      o = pcg_advance(o, 2u * (uint64_t)next);
      j0 += (uint32_t)np;
      __syncthreads();  // s_acc and s_cand are rewritten by the next round
    }
  }
  if (tid == 0) *grng_out = o;
}
