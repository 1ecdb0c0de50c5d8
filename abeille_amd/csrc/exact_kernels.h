// exact_kernels.h - ExactMGCancelator / BasicExactMGCancelator
// (src/exact_mg_cancelator.cpp, src/basic_exact_mg_cancelator.cpp) on a
// device-resident fission bank and its parent records.
//
// Pipeline (the host drives it, abx.hip):
//   k_ex_keys      one lane per site: region number (exact.h) << 32 | site
//   radix sort     regions in order, each region's sites in bank order
//   k_ex_heads     region starts; an exclusive scan numbers the regions
//   k_ex_bins      one record per region: its range of sorted positions
//   k_ex_bin_init  one wave per region: signs present, W and W2 in bank order
//   k_ex_samples   one lane per region: the n_samples points (Sobol or pcg32)
//   k_ex_f         one lane per site: <f> and <1/f> over the points
//   k_ex_cancel    one wave per region: the C sums, then cancel_bin for wgt
//                  and wgt2 with their uniform weights, in bank order
//   k_ex_new       one block: the uniform particles, in region order, from
//                  the one global pcg32 stream
// Every floating-point sum the reference forms over a region's particles is
// formed here in the same order (a wave reads 64 values, then adds them one
// by one), so the weights and new particles are bit-identical to a serial
// pass over the regions in order.
#pragma once

#include "exact.h"
#include "kernels.h"

struct ExBinD {
  int64_t start, end;  // sorted positions [start, end)
  uint64_t reg;        // region number
  int32_t m1, m2;      // wgt / wgt2 of both signs present
  int32_t avg;         // <f> and <1/f> are computed for it
  int32_t can_cancel;
  double W, W2;
  double sum_c, sum_c_wgt, sum_c_wgt2;
  double uw, uw2;      // uniform weights
  uint32_t n_new;      // uniform particles
  int32_t pad_;
};

struct ExSample {
  double x, y, z;
  int32_t g;
  int32_t pad_;
};

__device__ __forceinline__ uint64_t ex_reg_of(uint64_t key) { return key >> 32; }
__device__ __forceinline__ uint32_t ex_site_of(uint64_t key) { return (uint32_t)key; }

// Adds lanes 0..cnt-1 of v to acc, in lane order (all lanes get the sum).
__device__ __forceinline__ double ex_wave_add(double acc, double v, int cnt) {
  for (int l = 0; l < cnt; l++) acc += __shfl(v, l, 64);
  return acc;
}

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

__global__ void k_ex_bins(const uint64_t* __restrict__ keys, int64_t n,
                          const uint32_t* __restrict__ head, const uint64_t* __restrict__ offs,
                          ExBinD* __restrict__ bins) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t r = ex_reg_of(keys[t]);
  if (r == kExNone) return;
  const uint64_t b = offs[t] + head[t] - 1u;
  if (head[t]) {
    bins[b].start = t;
    bins[b].reg = r;
  }
  if (t == n - 1 || ex_reg_of(keys[t + 1]) != r) bins[b].end = t + 1;
}

// add_particle's sums and mixed-sign tests of each region
__global__ void k_ex_bin_init(DevExact E, const abx_banked* __restrict__ bank,
                              const uint64_t* __restrict__ keys, ExBinD* __restrict__ bins,
                              int64_t nb) {
  const int lane = threadIdx.x & 63;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= nb) return;
  const int64_t s0 = bins[b].start, s1 = bins[b].end;
  bool p1 = false, n1 = false, p2 = false, n2 = false;
  double W = 0., W2 = 0.;
  for (int64_t c = s0; c < s1; c += 64) {
    const int64_t t = c + lane;
    const int cnt = (int)(s1 - c < 64 ? s1 - c : 64);
    double w = 0., w2 = 0.;
    if (t < s1) {
      const abx_banked& q = bank[ex_site_of(keys[t])];
      w = q.wgt;
      w2 = q.wgt2;
      if (w > 0.)
        p1 = true;
      else if (w < 0.)
        n1 = true;
      if (w2 > 0.)
        p2 = true;
      else if (w2 < 0.)
        n2 = true;
    }
    W = ex_wave_add(W, w, cnt);
    W2 = ex_wave_add(W2, w2, cnt);
  }
  const bool m1 = __ballot(p1) != 0 && __ballot(n1) != 0;
  const bool m2 = __ballot(p2) != 0 && __ballot(n2) != 0;
  if (lane == 0) {
    ExBinD& B = bins[b];
    B.m1 = m1;
    B.m2 = m2;
    const bool needs = E.kind == ABX_EXACT_MG || E.beta == ABX_BETA_AVERAGE_F ||
                       E.beta == ABX_BETA_AVERAGE_G;
    B.avg = (s1 - s0 > 1) && (m1 || m2) && needs;
    B.can_cancel = 1;
    B.W = W;
    B.W2 = W2;
    B.sum_c = B.sum_c_wgt = B.sum_c_wgt2 = 0.;
    B.uw = B.uw2 = 0.;
    B.n_new = 0;
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
// :329-341; get_averages :262-283).  The exact cancelator gives up on the
// region at the first f of zero.
__global__ void k_ex_f(DevProblem P, DevExact E, const abx_banked* __restrict__ bank,
                       const abx_parent_info* __restrict__ par, const uint64_t* __restrict__ keys,
                       int64_t n, const uint32_t* __restrict__ head,
                       const uint64_t* __restrict__ offs, ExBinD* __restrict__ bins,
                       const ExSample* __restrict__ smp, double* __restrict__ fv,
                       double* __restrict__ fi) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t key = keys[t];
  if (ex_reg_of(key) == kExNone) return;
  const uint64_t b = offs[t] + head[t] - 1u;
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
  fv[t] = sf / (double)E.n_samples;
  fi[t] = sfi / (double)E.n_samples;
}

// The C sums (compute_averages :343-364), then cancel_bin for wgt and for
// wgt2 (exact :367-403, basic :438-468): beta, the site's own f, the kept
// fraction P_p and the uniform share P_u.  A non-finite P_p or P_u, or (basic)
// a zero weight, ends the region's pass, as the reference's `return` does.
__global__ void k_ex_cancel(DevProblem P, DevExact E, abx_banked* __restrict__ bank,
                            const abx_parent_info* __restrict__ par,
                            const uint64_t* __restrict__ keys, ExBinD* __restrict__ bins,
                            int64_t nb, const double* __restrict__ fv,
                            const double* __restrict__ fi) {
  const int lane = threadIdx.x & 63;
  const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= nb) return;
  const ExBinD B = bins[b];
  const int64_t s0 = B.start, s1 = B.end;
  if (s1 - s0 <= 1) return;
  const bool basic = E.kind == ABX_BASIC_EXACT_MG;
  const bool can = B.can_cancel != 0;
  double sum_c = 0., scw = 0., scw2 = 0.;
  if (B.avg && can && (!basic || E.beta == ABX_BETA_AVERAGE_G)) {
    for (int64_t c = s0; c < s1; c += 64) {
      const int64_t t = c + lane;
      const int cnt = (int)(s1 - c < 64 ? s1 - c : 64);
      double cc = 0., cw = 0., cw2 = 0.;
      if (t < s1) {
        const abx_banked& q = bank[ex_site_of(keys[t])];
        cc = ex_c(fv[t], fi[t]);
        cw = cc * q.wgt;
        cw2 = cc * q.wgt2;
      }
      sum_c = ex_wave_add(sum_c, cc, cnt);
      scw = ex_wave_add(scw, cw, cnt);
      scw2 = ex_wave_add(scw2, cw2, cnt);
    }
  }
  const ExKeyD K = ex_decode(E, B.reg);
  double uw[2] = {0., 0.};
  for (int pass = 0; pass < 2; pass++) {
    const bool first = pass == 0;
    if (!(first ? B.m1 : B.m2)) continue;
    double acc = 0.;
    for (int64_t c = s0; c < s1; c += 64) {
      const int64_t t = c + lane;
      const int cnt = (int)(s1 - c < 64 ? s1 - c : 64);
      bool stop = false;
      double wgt = 0., Pp = 0., Pu = 0.;
      uint32_t site = 0;
      if (t < s1) {
        site = ex_site_of(keys[t]);
        const abx_banked s = bank[site];
        const abx_parent_info q = par[site];
        const V3 r1{q.r1[0], q.r1[1], q.r1[2]};
        const V3 r4{s.r[0], s.r[1], s.r[2]};
        wgt = first ? s.wgt : s.wgt2;
        if (basic && wgt == 0.) {
          stop = true;
        } else {
          double Bt = 0.;  // get_beta
          if (can) {
            if (!basic || E.beta == ABX_BETA_AVERAGE_G) {
              const double S = (first ? scw : scw2) / (1. + sum_c);
              const double f = fv[t], f_inv = fi[t];
              Bt = f * (1. / (2. * f * f_inv - 1.)) * (1. - (S / wgt));
            } else if (E.beta == ABX_BETA_MINIMUM) {
              Bt = ex_basic_min_f(E, K, r1, q.Esmp);
            } else if (E.beta == ABX_BETA_AVERAGE_F) {
              const double N = (double)(s1 - s0);
              const double W = first ? B.W : B.W2;
              Bt = fv[t] * (1. - ((W) / ((N + 1.) * wgt)));
            }
          }
          double f;
          if (basic) {
            f = ex_basic_f(r4, r1, q.Esmp);
          } else {
            const V3 u1{q.u1[0], q.u1[1], q.u1[2]};
            f = ex_exact_f(E, P, K.mat, r1, u1, ex_group(P, q.E1), ex_group(P, q.E3), r4,
                           ex_group(P, s.E), q.Esmp);
          }
          Pp = (f - Bt) / f;
          Pu = Bt / f;
          if (isinf(Pu) || isinf(Pp) || isnan(Pu) || isnan(Pp)) stop = true;
        }
      }
      const uint64_t smask = __ballot(t < s1 && stop);
      const int upto = smask ? (int)__ffsll((long long)smask) - 1 : cnt;
      acc = ex_wave_add(acc, wgt * Pu, upto);
      if (lane < upto) {
        if (first)
          bank[site].wgt = wgt * Pp;
        else
          bank[site].wgt2 = wgt * Pp;
      }
      if (smask) break;
    }
    uw[pass] = acc;
  }
  if (lane == 0) {
    ExBinD& O = bins[b];
    O.sum_c = sum_c;
    O.sum_c_wgt = scw;
    O.sum_c_wgt2 = scw2;
    O.uw = uw[0];
    O.uw2 = uw[1];
    const double a = fabs(uw[0]), a2 = fabs(uw[1]);
    O.n_new = (uint32_t)ceil(a < a2 ? a2 : a);
  }
}

__global__ void k_ex_counts(const ExBinD* __restrict__ bins, int64_t nb, uint32_t* __restrict__ c) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) c[b] = bins[b].n_new;
}

// get_new_particles (exact :519-605, basic :564-611): each region's uniform
// particles in region order from the one pcg32 stream, each at the first of
// at most 100 candidate points (three numbers each) in the region's
// material.  One block: its lanes take the next particles assuming each
// accepts its first candidate (14 or 12 pcg32 steps apart); the particles
// before the first lane whose candidate fails are final.  That particle's
// later candidates are then tried side by side, the first that lands in the
// material wins, and the stream continues after it.
__global__ void __launch_bounds__(256) k_ex_new(DevProblem P, DevExact E,
                                                const ExBinD* __restrict__ bins, int64_t nb,
                                                const uint64_t* __restrict__ first_new,
                                                uint64_t grng, abx_banked* __restrict__ out,
                                                uint64_t* __restrict__ grng_out, int32_t* err) {
  __shared__ int s_min;
  TrkT<true> t;
  bind_lds_cursor(t, P);
  const int tid = threadIdx.x, nt = blockDim.x;
  const uint64_t tail = ex_tail_steps(E);
  const uint64_t pstep = 6u + tail;  // a particle whose first candidate is accepted
  uint64_t o = grng;                 // every lane keeps the same copy
  for (int64_t b = 0; b < nb; b++) {
    const ExBinD B = bins[b];
    const uint32_t N = B.n_new;
    if (N == 0) continue;
    const double w = B.uw / (double)N, w2 = B.uw2 / (double)N;
    const ExKeyD K = ex_decode(E, B.reg);
    abx_banked* dst = out + first_new[b];
    uint32_t j0 = 0;
    while (j0 < N) {
      const uint32_t cnt = N - j0 < (uint32_t)nt ? N - j0 : (uint32_t)nt;
      uint64_t s = 0;
      V3 r{0., 0., 0.};
      bool acc = true;
      if ((uint32_t)tid < cnt) {
        s = pcg_advance(o, (uint64_t)tid * pstep);
        const double x0 = rng_rand(s), x1 = rng_rand(s), x2 = rng_rand(s);
        r = ex_point(E, K, x0, x1, x2);
        acc = ex_material(P, t, r) == K.mat;
      }
      if (tid == 0) s_min = nt;
      __syncthreads();
      if (!acc) atomicMin(&s_min, tid);
      __syncthreads();
      const uint32_t firstbad = (uint32_t)s_min;
      __syncthreads();
      const uint32_t done = firstbad < cnt ? firstbad : cnt;
      if ((uint32_t)tid < done) {
        abx_banked u{};
        u.r[0] = r.x;
        u.r[1] = r.y;
        u.r[2] = r.z;
        ex_new_tail(E, P, K, s, u);
        u.wgt = w;
        u.wgt2 = w2;
        dst[j0 + tid] = u;
      }
      o = pcg_advance(o, (uint64_t)done * pstep);
      j0 += done;
      if (done == cnt) continue;
      // particle j0 failed its first candidate: candidates 1..99 side by side
      bool hit = false;
      if (tid >= 1 && tid < 100) {
        s = pcg_advance(o, (uint64_t)tid * 6u);
        const double x0 = rng_rand(s), x1 = rng_rand(s), x2 = rng_rand(s);
        r = ex_point(E, K, x0, x1, x2);
        hit = ex_material(P, t, r) == K.mat;
      }
      if (tid == 0) s_min = 100;
      __syncthreads();
      if (hit) atomicMin(&s_min, tid);
      __syncthreads();
      const int k = s_min;
      __syncthreads();
      if (k >= 100) {  // "Couldn't sample position for uniform particle."
        if (tid == 0) atomicCAS(err, 0, ABX_E_INVALID);
        return;
      }
      if (tid == k) {
        abx_banked u{};
        u.r[0] = r.x;
        u.r[1] = r.y;
        u.r[2] = r.z;
        ex_new_tail(E, P, K, s, u);
        u.wgt = w;
        u.wgt2 = w2;
        dst[j0] = u;
      }
      o = pcg_advance(o, (uint64_t)(k + 1) * 6u + tail);
      j0 += 1;
    }
  }
  if (tid == 0) *grng_out = o;
}
