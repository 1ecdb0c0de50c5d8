// hd_host.cpp - runs the device transport code (abeille_amd/csrc/physics.h,
// compiled for the host CPU) one history at a time, serially, so the CPU test
// suite can compare its per-history traces with the CPU restatement before any
// GPU is involved.  Test infrastructure: the product never builds or loads it.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../abeille_amd/csrc/exact.h"
#include "../../abeille_amd/csrc/source.h"
#include "../../abeille_amd/csrc/tables.h"

struct HostSink {
  static constexpr bool split = false;
  static constexpr bool ext = true;  // keeps the parent records
  std::vector<SecRec> stack;
  std::vector<abx_banked>* bank;
  std::vector<abx_parent_info>* parents;
  std::vector<abx_trace_rec>* trace_buf;
  int trace_cap = 0;
  int32_t n_rec = 0;
  int err = 0;
  uint64_t err_hist = 0;
  abx_diag diag{};
  void tally_add(double* a, double v) {
    *a += v;
    diag.tally_updates++;
  }
  bool sec_push(const SecRec& q) {
    if (stack.size() >= ABX_SEC_CAP) return false;
    stack.push_back(q);
    return true;
  }
  bool sec_pop(SecRec& q) {
    if (stack.empty()) return false;
    q = stack.back();
    stack.pop_back();
    return true;
  }
  void bank_fission(const Part& p, const V3& u, double E, double w) {
    abx_banked b{};
    b.r[0] = p.r.x; b.r[1] = p.r.y; b.r[2] = p.r.z;
    b.u[0] = u.x; b.u[1] = u.y; b.u[2] = u.z;
    b.E = E; b.wgt = w; b.wgt2 = 0.;
    b.parent_history_id = p.hid;
    b.parent_daughter_id = p.daughters;
    b.family_id = p.fam;
    bank->push_back(b);
    abx_parent_info q{};
    q.r1[0] = p.pr.x; q.r1[1] = p.pr.y; q.r1[2] = p.pr.z;
    q.u1[0] = p.pu.x; q.u1[1] = p.pu.y; q.u1[2] = p.pu.z;
    q.E1 = p.pE; q.E3 = p.E; q.Esmp = p.Esmp; q.prev_virtual = p.pvirt;
    parents->push_back(q);
  }
  template <class TK>
  void trace_rec(const Part& p, const TK& t, int ev, const DevProblem& P) {
    if (!trace_buf) return;
    if (n_rec < trace_cap) {
      abx_trace_rec r{};
      r.r[0] = p.r.x; r.r[1] = p.r.y; r.r[2] = p.r.z;
      r.u[0] = p.u.x; r.u[1] = p.u.y; r.u[2] = p.u.z;
      r.wgt = p.w;
      r.history_id = p.hid;
      r.secondary_id = p.sec_id;
      r.event = ev;
      r.cell = t.cell;
      r.material = t.cell >= 0 ? P.cell[t.cell].fill : -1;
      r.group = p.g;
      r.rng_calls = (int32_t)p.calls;
      trace_buf->push_back(r);
    }
    n_rec++;
  }
  const DevProblem* Pp = nullptr;
  template <class TK>
  void trace(const Part& p, const TK& t, int ev) { trace_rec(p, t, ev, *Pp); }
  void error(int code, uint64_t h) {
    if (!err) { err = code; err_hist = h; }
  }
  void count_flight() { diag.flights++; }
  void count_crossing() { diag.crossings++; }
  void count_collision() { diag.collisions++; }
  void count_virtual() { diag.virtual_collisions++; }
  void count_lost() { diag.lost++; }
  void count_secondary() { diag.secondaries++; }
};

struct hd_problem {
  HostTables T;
  std::vector<double> tally;
  DevProblem P;
  std::vector<double> src_cdf;
  DevSources S;
  std::vector<abx_parent_info> parents;  // one per banked site of the last transport
};

static std::string g_err;

extern "C" {
const char* hd_last_error() { return g_err.c_str(); }

int hd_create(const abx_problem_desc* d, hd_problem** out) {
  try {
    auto* h = new hd_problem;
    build_tables(d, h->T);
    h->tally.assign((size_t)std::max<int64_t>(1, h->T.tally_bins), 0.);
    auto map = [](const auto& v) { return v.data(); };
    h->P = dev_view(h->T, map, h->tally.data());
    std::vector<double> w;
    for (auto& s : h->T.sources) w.push_back(s.weight);
    h->src_cdf.resize(w.size());
    build_discrete_cdf(w.data(), (int)w.size(), h->src_cdf.data());
    h->S.src = h->T.sources.data();
    h->S.cdf = h->src_cdf.data();
    h->S.n = (int32_t)w.size();
    *out = h;
    return 0;
  } catch (const TableError& e) {
    g_err = e.what();
    return e.code;
  }
}

void hd_destroy(hd_problem* h) { delete h; }

struct HostCursor {
  std::vector<unsigned char> buf;
  TrkT<true> t;
  explicit HostCursor(const DevProblem& P) : buf((size_t)trk_bytes(P.trk_D, P.trk_L) + 16) {
    trk_bind(t, buf.data(), 1, 0, P.trk_D, P.trk_L);
  }
};

int hd_sample_sources(hd_problem* h, uint64_t first, size_t n, abx_particle* out) {
  HostCursor c(h->P);
  for (size_t i = 0; i < n; i++) {
    int rc = sample_source(h->P, h->S, first + i, c.t, out[i]);
    if (rc) return rc;
  }
  return 0;
}

int hd_transport(hd_problem* h, const abx_particle* bank, size_t n, double kcol,
                 int converged, abx_banked* fis_out, size_t cap, size_t* n_out,
                 abx_scores* scores, abx_diag* diag, abx_trace_rec* trace,
                 int trace_cap, int32_t* n_rec) {
  std::vector<abx_banked> fb;
  h->parents.clear();
  Scores6 sc{0, 0, 0, 0, 0, 0};
  abx_diag dg{};
  for (size_t i = 0; i < n; i++) {
    HostSink s;
    s.Pp = &h->P;
    s.bank = &fb;
    s.parents = &h->parents;
    std::vector<abx_trace_rec> tr;
    s.trace_buf = trace ? &tr : nullptr;
    s.trace_cap = trace_cap;
    Part p{};
    const abx_particle& a = bank[i];
    p.r = V3{a.r[0], a.r[1], a.r[2]};
    p.u = V3{a.u[0], a.u[1], a.u[2]};
    p.rb = p.r;
    p.E = a.E;
    p.g = energy_index(h->P, a.E);
    p.w = a.wgt;
    p.w2 = a.wgt2;
    p.rng = a.rng_state;
    p.hid = a.history_id;
    p.fam = a.family_id;
    p.alive = 1;
    p.pu = V3{1., 0., 0.};
    HostCursor cur(h->P);
    switch (h->P.tracking) {
      case 0: run_history<0>(h->P, s, p, cur.t, kcol, converged != 0, sc); break;
      case 1: run_history<1>(h->P, s, p, cur.t, kcol, converged != 0, sc); break;
      default: run_history<2>(h->P, s, p, cur.t, kcol, converged != 0, sc); break;
    }
    if (s.err) {
      g_err = "history " + std::to_string(s.err_hist) + " failed";
      return s.err;
    }
    dg.histories++;
    dg.flights += s.diag.flights;
    dg.crossings += s.diag.crossings;
    dg.collisions += s.diag.collisions;
    dg.virtual_collisions += s.diag.virtual_collisions;
    dg.lost += s.diag.lost;
    dg.secondaries += s.diag.secondaries;
    dg.tally_updates += s.diag.tally_updates;
    if (trace) {
      std::copy(tr.begin(), tr.end(), trace + i * (size_t)trace_cap);
      n_rec[i] = s.n_rec;
    }
  }
  *n_out = fb.size();
  if (fb.size() > cap) return ABX_E_CAPACITY;
  std::copy(fb.begin(), fb.end(), fis_out);
  scores->k_col = sc.k_col; scores->k_abs = sc.k_abs; scores->k_trk = sc.k_trk;
  scores->k_tot = sc.k_tot; scores->leakage = sc.leak; scores->mig = sc.mig;
  *diag = dg;
  return 0;
}

int hd_parents_fetch(hd_problem* h, abx_parent_info* out, size_t cap, size_t* n) {
  *n = h->parents.size();
  if (h->parents.size() > cap) return ABX_E_CAPACITY;
  std::copy(h->parents.begin(), h->parents.end(), out);
  return 0;
}

int hd_tally_fetch(hd_problem* h, double* out) {
  std::copy(h->tally.begin(), h->tally.end(), out);
  return 0;
}
}

// ---- RNG / math probes of the device headers (rng.h, detmath.h) ----------
extern "C" {
void hd_rng_stream(uint64_t seed, uint64_t stride, uint64_t h, int n, uint32_t* out) {
  uint64_t s = pcg_history_state(seed, stride, h);
  for (int i = 0; i < n; i++) out[i] = pcg_next(s);
}
void hd_rng_rand(uint64_t seed, uint64_t stride, uint64_t h, int n, double* out) {
  uint64_t s = pcg_history_state(seed, stride, h);
  for (int i = 0; i < n; i++) out[i] = rng_rand(s);
}
void hd_rng_exponential(uint64_t seed, uint64_t stride, uint64_t h, int n, double lambda,
                        double* out) {
  uint64_t s = pcg_history_state(seed, stride, h);
  for (int i = 0; i < n; i++) out[i] = rng_exponential(s, lambda);
}
// n draws of discrete(w) with the cdf built once; returns the next raw output
uint32_t hd_rng_discrete(uint64_t seed, uint64_t stride, uint64_t h, const double* w, int nw,
                         int n, int32_t* out) {
  std::vector<double> cdf(nw);
  build_discrete_cdf(w, nw, cdf.data());
  uint64_t s = pcg_history_state(seed, stride, h);
  for (int i = 0; i < n; i++) out[i] = rng_discrete(s, cdf.data(), nw);
  return pcg_next(s);
}
uint32_t hd_pcg_seeded_stream(uint64_t state_seed, uint64_t stream, int n, uint32_t* out) {
  // pcg32(state, stream): inc = (stream << 1) | 1, state = bump(seed + inc)
  const uint64_t inc = (stream << 1) | 1u;
  uint64_t s = (state_seed + inc) * ABX_PCG_MULT + inc;
  for (int i = 0; i < n; i++) {
    const uint64_t old = s;
    s = old * ABX_PCG_MULT + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    out[i] = (xs >> rot) | (xs << ((0u - rot) & 31u));
  }
  return 0;
}
double hd_log(double x) { return dm_log(x); }
double hd_sobol(uint64_t index, int dim) {
  static uint64_t M[kSobolDims * kSobolBits];
  static bool built = false;
  if (!built) {
    sobol_build(M);
    built = true;
  }
  return sobol_at(M, index, dim);
}
double hd_exp(double x) { return dm_exp(x); }
void hd_sincos(double x, double* s, double* c) { dm_sincos(x, s, c); }
}
