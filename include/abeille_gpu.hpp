// abeille_gpu.hpp - C++ host mirror of the reference's transport interface
// on top of the C ABI (abx.h).
//
// Same names, argument meaning and error behaviour as the reference's
//   Transporter          include/simulation/transporter.hpp:39-87
//   Tallies              include/simulation/tallies.hpp:38-186
//   Particle / BankedParticle  include/simulation/particle.hpp:38-242
//   Cancelator           include/simulation/cancelator.hpp:36-45
// so a driver written against them (FixedSource::run, PowerIterator::run)
// runs unchanged with GpuTransporter in place of DeltaTracker /
// SurfaceTracker / CarterTracker, and GpuExactCancelator in place of
// ExactMGCancelator / BasicExactMGCancelator.  Fatal conditions throw abx::Error (the
// reference calls fatal_error(), which exits).
#ifndef ABEILLE_GPU_HPP
#define ABEILLE_GPU_HPP

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "abx.h"

namespace abx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != ABX_OK) throw Error(rc, abx_last_error());
}

struct Vec3 {
  double x = 0, y = 0, z = 0;
};

// BankedParticle (particle.hpp:38-66).  The exact-cancellation fields are
// filled when the problem has an exact cancelator.
struct BankedParticle {
  Vec3 r, u;
  double E = 0, wgt = 0, wgt2 = 0;
  uint64_t parent_history_id = 0, parent_daughter_id = 0, family_id = 0;
  bool parents_previous_was_virtual = false;
  Vec3 parents_previous_position;
  Vec3 parents_previous_direction{1., 0., 0.};
  double parents_previous_previous_energy = 0;
  double parents_previous_energy = 0;
  double Esmp_parent = 0.;
};

// pcg32 (pcg-cpp, default stream) as far as the drivers use it: seeded like
// settings::rng, its state handed to the device where the cancellation draws.
struct pcg32 {
  uint64_t state = 0;
  explicit pcg32(uint64_t seed = 0) { this->seed(seed); }
  void seed(uint64_t s) {
    const uint64_t M = 6364136223846793005ULL, I = 1442695040888963407ULL;
    state = (s + I) * M + I;
  }
};

// Particle as handed to transport(): state plus its pcg32 stream.
class Particle {
 public:
  Particle(Vec3 r, Vec3 u, double E, double wgt, uint64_t id = 0)
      : r_(r), u_(u), E_(E), w_(wgt), w2_(0.), id_(id) {}
  Particle(Vec3 r, Vec3 u, double E, double wgt, double wgt2, uint64_t id = 0)
      : r_(r), u_(u), E_(E), w_(wgt), w2_(wgt2), id_(id) {}
  const Vec3& r() const { return r_; }
  const Vec3& u() const { return u_; }
  double E() const { return E_; }
  double wgt() const { return w_; }
  double wgt2() const { return w2_; }
  uint64_t history_id() const { return id_; }
  uint64_t family_id() const { return fam_; }
  void set_family_id(uint64_t f) { fam_ = f; }
  // Particle::initialize_rng (particle.hpp:188-193): pcg32(seed) advanced
  // by stride * history_id (Brown's jump).
  void initialize_rng(uint64_t seed, uint64_t stride) {
    const uint64_t M = 6364136223846793005ULL, I = 1442695040888963407ULL;
    uint64_t s = (seed + I) * M + I;
    uint64_t delta = stride * id_, cm = M, cp = I, am = 1, ap = 0;
    while (delta) {
      if (delta & 1u) {
        am *= cm;
        ap = ap * cm + cp;
      }
      cp = (cm + 1) * cp;
      cm *= cm;
      delta >>= 1;
    }
    rng_state = am * s + ap;
  }
  uint64_t rng_state = 0;

 private:
  Vec3 r_, u_;
  double E_, w_, w2_;
  uint64_t id_, fam_ = 0;
};

// Owns the staged problem (abx_create) - the object graph that
// parse_input_file() builds in the reference.
class Problem {
 public:
  explicit Problem(const abx_problem_desc& d) : exact_(d.exact_cancel != nullptr) {
    check(abx_create(&d, &p_));
  }
  ~Problem() { abx_destroy(p_); }
  Problem(const Problem&) = delete;
  Problem& operator=(const Problem&) = delete;
  abx_problem* get() const { return p_; }
  // the description names an exact cancelator: sites carry parent records
  bool exact_cancellation() const { return exact_; }

 private:
  abx_problem* p_ = nullptr;
  bool exact_ = false;
};

// settings::converged (src/settings.cpp): set by the driver once the
// inactive generations are done; transport() reads it to gate the mesh
// tallies, as the reference's trackers pass it to Tallies::score_collision /
// score_flight.
namespace settings {
inline bool converged = false;
}

// Tallies (tallies.hpp:38-186): scalar scores on the host, mesh tallies
// resident on the device (their generation arrays, means and variances).
class Tallies {
 public:
  Tallies(std::shared_ptr<Problem> prob, double tot_wgt) : prob_(std::move(prob)), W_(tot_wgt) {}
  void score_k_col(double s) { k_col_s_ += s; }
  void score_k_abs(double s) { k_abs_s_ += s; }
  void score_k_trk(double s) { k_trk_s_ += s; }
  void score_k_tot(double s) { k_tot_s_ += s; }
  void score_leak(double s) { leak_s_ += s; }
  void score_mig_area(double s) { mig_s_ += s; }
  // Tallies::clear_generation (tallies.cpp:141-157)
  void clear_generation() {
    k_col_s_ = k_abs_s_ = k_trk_s_ = k_tot_s_ = leak_s_ = mig_s_ = 0.;
    check(abx_tally_clear(prob_->get()));
  }
  // Tallies::calc_gen_values (tallies.cpp:159-181)
  void calc_gen_values() {
    k_col_ = k_col_s_ / W_;
    k_abs_ = k_abs_s_ / W_;
    k_trk_ = k_trk_s_ / W_;
    leak_ = leak_s_ / W_;
    k_tot_ = k_tot_s_ / W_;
    mig_ = mig_s_ / W_;
    k_col_vec.push_back(k_col_);
    k_abs_vec.push_back(k_abs_);
    k_trk_vec.push_back(k_trk_);
    leak_vec.push_back(leak_);
    mig_vec.push_back(mig_);
  }
  // Tallies::record_generation (tallies.cpp:183-203): the scalars here, the
  // mesh tallies on the device (MeshTally::record_generation).
  void record_generation(double multiplier = 1.) {
    gen_++;
    upd(k_col_, k_col_avg_, k_col_var_);
    upd(k_abs_, k_abs_avg_, k_abs_var_);
    upd(k_trk_, k_trk_avg_, k_trk_var_);
    upd(leak_, leak_avg_, leak_var_);
    upd(k_tot_, k_tot_avg_, k_tot_var_);
    upd(mig_, mig_avg_, mig_var_);
    check(abx_tally_record(prob_->get(), multiplier));
  }
  void set_total_weight(double W) { W_ = W; }
  double total_weight() const { return W_; }

  double kcol() const { return k_col_; }
  double kcol_avg() const { return k_col_avg_; }
  double kcol_err() const { return err(k_col_var_); }
  double kabs() const { return k_abs_; }
  double kabs_avg() const { return k_abs_avg_; }
  double kabs_err() const { return err(k_abs_var_); }
  double ktrk() const { return k_trk_; }
  double ktrk_avg() const { return k_trk_avg_; }
  double ktrk_err() const { return err(k_trk_var_); }
  double ktot() const { return k_tot_; }
  double ktot_avg() const { return k_tot_avg_; }
  double ktot_err() const { return err(k_tot_var_); }
  double leakage() const { return leak_; }
  double leakage_avg() const { return leak_avg_; }
  double leakage_err() const { return err(leak_var_); }
  double mig_area() const { return mig_; }
  double mig_area_avg() const { return mig_avg_; }
  double mig_area_err() const { return err(mig_var_); }
  int generations() const { return gen_; }
  // A mesh tally's mean and error of the mean, [Ne][Nx][Ny][Nz]
  // (MeshTally::write_tally, mesh_tally.cpp:191-197).
  void mesh_tally_stats(int tally, std::vector<double>& avg, std::vector<double>& err) const {
    size_t nb = 0;
    check(abx_tally_size(prob_->get(), tally, &nb));
    avg.resize(nb);
    err.resize(nb);
    check(abx_tally_stats(prob_->get(), tally, avg.data(), err.data()));
  }
  std::vector<double> k_col_vec, k_abs_vec, k_trk_vec, leak_vec, mig_vec;
  Problem& problem() { return *prob_; }

 private:
  // Tallies::update_avg_and_var (tallies.cpp:268-280)
  void upd(double x, double& avg, double& var) {
    const double dg = (double)gen_, old = avg, vold = var;
    avg = old + (x - old) / dg;
    if (gen_ > 1) var = vold + ((x - old) * (x - old) / dg) - (vold / (dg - 1.));
  }
  double err(double var) const { return std::sqrt(var / (double)gen_); }
  std::shared_ptr<Problem> prob_;
  double W_;
  double k_col_s_ = 0, k_abs_s_ = 0, k_trk_s_ = 0, k_tot_s_ = 0, leak_s_ = 0, mig_s_ = 0;
  double k_col_ = 1.0, k_abs_ = 1.0, k_trk_ = 1.0, k_tot_ = 1.0, leak_ = 0., mig_ = 0.;
  double k_col_avg_ = 0, k_col_var_ = 0, k_abs_avg_ = 0, k_abs_var_ = 0, k_trk_avg_ = 0,
         k_trk_var_ = 0, k_tot_avg_ = 0, k_tot_var_ = 0, leak_avg_ = 0, leak_var_ = 0,
         mig_avg_ = 0, mig_var_ = 0;
  int gen_ = 0;
};

// The noise source of the reference's noise transport (out of this path:
// transport() with noise == true is refused).
class NoiseMaker;

// Transporter (transporter.hpp:39-87).
class Transporter {
 public:
  explicit Transporter(std::shared_ptr<Tallies> t) : tallies(std::move(t)) {}
  virtual ~Transporter() = default;
  virtual std::vector<BankedParticle> transport(std::vector<Particle>& bank, bool noise = false,
                                                std::vector<BankedParticle>* noise_bank = nullptr,
                                                const NoiseMaker* noise_maker = nullptr) = 0;

 protected:
  std::shared_ptr<Tallies> tallies;
};

// The GPU transporter: whichever tracker the problem's settings name
// (delta / surface / carter) runs on the device.  `bank` is consumed, the
// fission bank comes back in parent order then push order, the scores go
// to the Tallies, and the mesh tallies score when settings::converged.
class GpuTransporter : public Transporter {
 public:
  using Transporter::Transporter;
  std::vector<BankedParticle> transport(std::vector<Particle>& bank, bool noise = false,
                                        std::vector<BankedParticle>* noise_bank = nullptr,
                                        const NoiseMaker* noise_maker = nullptr) override {
    (void)noise_bank;
    (void)noise_maker;
    if (noise) throw Error(ABX_E_INVALID, "noise transport is not part of the GPU path");
    std::vector<abx_particle> in(bank.size());
    for (size_t i = 0; i < bank.size(); i++) {
      const Particle& p = bank[i];
      abx_particle& a = in[i];
      a.r[0] = p.r().x;
      a.r[1] = p.r().y;
      a.r[2] = p.r().z;
      a.u[0] = p.u().x;
      a.u[1] = p.u().y;
      a.u[2] = p.u().z;
      a.E = p.E();
      a.wgt = p.wgt();
      a.wgt2 = p.wgt2();
      a.history_id = p.history_id();
      a.family_id = p.family_id();
      a.rng_state = p.rng_state;
    }
    abx_problem* prob = tallies->problem().get();
    std::vector<abx_banked> out(2 * bank.size() + 1024);
    size_t n_out = 0;
    abx_scores sc{};
    abx_diag dg{};
    const int rc = abx_transport(prob, in.data(), in.size(), tallies->kcol(),
                                 settings::converged ? 1 : 0, out.data(), out.size(), &n_out,
                                 &sc, &dg);
    if (rc == ABX_E_CAPACITY && n_out > out.size()) {
      // more sites than the guess: the ordered bank waits on the device
      out.resize(n_out);
      check(abx_fission_fetch(prob, out.data(), out.size(), &n_out));
    } else {
      check(rc);
    }
    tallies->score_k_col(sc.k_col);
    tallies->score_k_abs(sc.k_abs);
    tallies->score_k_trk(sc.k_trk);
    tallies->score_k_tot(sc.k_tot);
    tallies->score_leak(sc.leakage);
    tallies->score_mig_area(sc.mig);
    std::vector<BankedParticle> res(n_out);
    for (size_t i = 0; i < n_out; i++) {
      const abx_banked& b = out[i];
      res[i].r = {b.r[0], b.r[1], b.r[2]};
      res[i].u = {b.u[0], b.u[1], b.u[2]};
      res[i].E = b.E;
      res[i].wgt = b.wgt;
      res[i].wgt2 = b.wgt2;
      res[i].parent_history_id = b.parent_history_id;
      res[i].parent_daughter_id = b.parent_daughter_id;
      res[i].family_id = b.family_id;
    }
    if (tallies->problem().exact_cancellation() && n_out) {
      // BankedParticle's cancellation fields (src/transporter.cpp:440-456)
      std::vector<abx_parent_info> par(n_out);
      size_t np = 0;
      check(abx_parent_fetch(prob, par.data(), par.size(), &np));
      for (size_t i = 0; i < n_out; i++) {
        const abx_parent_info& q = par[i];
        res[i].parents_previous_position = {q.r1[0], q.r1[1], q.r1[2]};
        res[i].parents_previous_direction = {q.u1[0], q.u1[1], q.u1[2]};
        res[i].parents_previous_previous_energy = q.E1;
        res[i].parents_previous_energy = q.E3;
        res[i].Esmp_parent = q.Esmp;
        res[i].parents_previous_was_virtual = q.prev_virtual != 0;
      }
    }
    bank.clear();
    return res;
  }
};

// Cancelator (cancelator.hpp:36-45).
class Cancelator {
 public:
  virtual ~Cancelator() = default;
  virtual bool add_particle(BankedParticle& p) = 0;
  virtual void perform_cancellation(pcg32& rng) = 0;
  virtual std::vector<BankedParticle> get_new_particles(pcg32& rng) = 0;
  virtual void clear() = 0;
};

// The exact or basic-exact cancelator the problem's description names, on
// the device.  add_particle keeps a pointer to the particle, as the
// reference's bins do; perform_cancellation runs the whole pass (regions,
// weights, uniform particles) from rng's state and writes the new weights
// through those pointers; get_new_particles hands over the uniform
// particles.  rng ends where the reference's settings::rng ends after both
// calls.  A particle outside every region is still kept and returns true
// (the device decides region membership; the reference's add_particle
// returns false for one outside the mesh, which only the message of
// perform_regional_cancellation counts).
class GpuExactCancelator : public Cancelator {
 public:
  explicit GpuExactCancelator(std::shared_ptr<Problem> prob) : prob_(std::move(prob)) {
    if (!prob_->exact_cancellation())
      throw Error(ABX_E_INVALID, "the problem's description names no exact cancelator");
  }
  bool add_particle(BankedParticle& p) override {
    parts_.push_back(&p);
    return true;
  }
  void perform_cancellation(pcg32& rng) override {
    const size_t n = parts_.size();
    std::vector<abx_banked> b(n);
    std::vector<abx_parent_info> q(n);
    for (size_t i = 0; i < n; i++) {
      const BankedParticle& p = *parts_[i];
      abx_banked& a = b[i];
      a.r[0] = p.r.x; a.r[1] = p.r.y; a.r[2] = p.r.z;
      a.u[0] = p.u.x; a.u[1] = p.u.y; a.u[2] = p.u.z;
      a.E = p.E; a.wgt = p.wgt; a.wgt2 = p.wgt2;
      a.parent_history_id = p.parent_history_id;
      a.parent_daughter_id = p.parent_daughter_id;
      a.family_id = p.family_id;
      abx_parent_info& r = q[i];
      r.r1[0] = p.parents_previous_position.x;
      r.r1[1] = p.parents_previous_position.y;
      r.r1[2] = p.parents_previous_position.z;
      r.u1[0] = p.parents_previous_direction.x;
      r.u1[1] = p.parents_previous_direction.y;
      r.u1[2] = p.parents_previous_direction.z;
      r.E1 = p.parents_previous_previous_energy;
      r.E3 = p.parents_previous_energy;
      r.Esmp = p.Esmp_parent;
      r.prev_virtual = p.parents_previous_was_virtual ? 1 : 0;
      r.pad_ = 0;
    }
    abx_problem* pr = prob_->get();
    check(abx_rng_set_state(pr, rng.state));
    std::vector<abx_banked> out(2 * n + 1024);
    size_t n_out = 0;
    int rc = abx_exact_cancel_bank(pr, b.data(), q.data(), n, out.data(), out.size(), &n_out);
    if (rc == ABX_E_CAPACITY && n_out > out.size()) {
      // more uniform particles than the guess: run the pass again from
      // the same rng state with room for them
      out.resize(n_out);
      check(abx_rng_set_state(pr, rng.state));
      rc = abx_exact_cancel_bank(pr, b.data(), q.data(), n, out.data(), out.size(), &n_out);
    }
    check(rc);
    check(abx_rng_get_state(pr, &rng.state));
    for (size_t i = 0; i < n; i++) {
      parts_[i]->wgt = out[i].wgt;
      parts_[i]->wgt2 = out[i].wgt2;
    }
    uniform_.clear();
    for (size_t i = n; i < n_out; i++) {
      BankedParticle u;
      u.r = {out[i].r[0], out[i].r[1], out[i].r[2]};
      u.u = {out[i].u[0], out[i].u[1], out[i].u[2]};
      u.E = out[i].E;
      u.wgt = out[i].wgt;
      u.wgt2 = out[i].wgt2;
      uniform_.push_back(u);
    }
  }
  std::vector<BankedParticle> get_new_particles(pcg32&) override {
    std::vector<BankedParticle> u;
    u.swap(uniform_);
    return u;
  }
  void clear() override {
    parts_.clear();
    uniform_.clear();
  }

 private:
  std::shared_ptr<Problem> prob_;
  std::vector<BankedParticle*> parts_;
  std::vector<BankedParticle> uniform_;
};

}  // namespace abx

#endif
