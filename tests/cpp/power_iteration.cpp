// power_iteration.cpp - the reference's PowerIterator::run loop
// (src/power_iterator.cpp:305-429) written against the C++ host mirror
// (include/abeille_gpu.hpp), with GpuTransporter as the Transporter.  Runs
// the LA-13511 PUa-1-0-SL critical slab (analytic k = 1) and prints k.
// Built and run by tests/test_cpp_host_api.py.
#include <cstdio>
#include <memory>
#include <vector>

#include "../../include/abeille_gpu.hpp"

int main(int argc, char** argv) {
  const int N = argc > 1 ? atoi(argv[1]) : 20000;
  const int ngen = argc > 2 ? atoi(argv[2]) : 60, nignored = argc > 3 ? atoi(argv[3]) : 10;
  // PUa-1-0-SL: one group, slab |x| < 1.853722 cm, vacuum in x, reflective y/z
  abx_surface s[6] = {};
  const double xs[6] = {-1.853722, 1.853722, -5, 5, -5, 5};
  for (int i = 0; i < 6; i++) {
    s[i].type = i / 2;  // x, y, z planes
    s[i].boundary = i < 2 ? ABX_VACUUM : ABX_REFLECTIVE;
    s[i].id = (uint32_t)(i + 1);
    s[i].p[0] = xs[i];
  }
  const int32_t rpn[6] = {1, -2, 3, -4, 5, -6};  // simple cell: intersection
  abx_cell cell{1, ABX_FILL_MATERIAL, 0, 0, 6};
  const int32_t ucells[1] = {0};
  abx_universe uni{};
  uni.id = 1;
  uni.type = ABX_CELL_UNIVERSE;
  uni.cells_offset = 0;
  uni.n_cells = 1;
  uni.outer = -1;
  uni.elems_offset = -1;
  const double tot[1] = {0.32640}, ab[1] = {0.101184}, sc[1] = {0.225216}, fi[1] = {0.081600},
               nu[1] = {3.24}, chi[1] = {1.0};
  abx_material mat{};
  mat.id = 1;
  mat.total = tot;
  mat.absorption = ab;
  mat.scatter = sc;
  mat.fission = fi;
  mat.nu_prompt = nu;
  mat.chi = chi;
  abx_source src{};
  src.spatial = ABX_SRC_POINT;
  src.direction = ABX_DIR_ISOTROPIC;
  src.energy = 0.5;
  src.weight = 1.0;
  const double eb[2] = {0., 1.};
  abx_problem_desc d{};
  d.n_surfaces = 6;
  d.surfaces = s;
  d.n_cells = 1;
  d.cells = &cell;
  d.n_rpn = 6;
  d.rpn = rpn;
  d.n_universes = 1;
  d.universes = &uni;
  d.n_universe_cells = 1;
  d.universe_cells = ucells;
  d.root_universe = 0;
  d.n_materials = 1;
  d.materials = &mat;
  d.n_sources = 1;
  d.sources = &src;
  d.settings.mode = ABX_K_EIGENVALUE;
  d.settings.tracking = ABX_SURFACE_TRACKING;
  d.settings.seed = 19073486328125ULL;
  d.settings.stride = 152917ULL;
  d.settings.wgt_cutoff = 0.25;
  d.settings.wgt_survival = 1.0;
  d.settings.wgt_split = 2.0;
  d.settings.nparticles = N;
  d.settings.ngroups = 1;
  d.settings.energy_bounds = eb;
  try {
    auto prob = std::make_shared<abx::Problem>(d);
    auto tallies = std::make_shared<abx::Tallies>(prob, (double)N);
    // PowerIterator::sample_source_from_sources
    std::vector<abx_particle> src_bank(N);
    abx::check(abx_sample_sources(prob->get(), 0, N, src_bank.data()));
    std::vector<abx::Particle> bank;
    for (auto& a : src_bank) {
      abx::Particle p({a.r[0], a.r[1], a.r[2]}, {a.u[0], a.u[1], a.u[2]}, a.E, a.wgt, a.history_id);
      p.rng_state = a.rng_state;
      p.set_family_id(a.history_id);
      bank.push_back(p);
    }
    uint64_t histories_counter = (uint64_t)N;
    std::shared_ptr<abx::Transporter> transporter =
        std::make_shared<abx::GpuTransporter>(tallies);
    // PowerIterator::run (src/power_iterator.cpp:305-429), without the
    // entropy, family, MPI and output bookkeeping
    if (nignored == 0) abx::settings::converged = true;
    for (int g = 1; g <= ngen; g++) {
      std::vector<abx::BankedParticle> next_gen = transporter->transport(bank);
      if (next_gen.size() == 0) throw abx::Error(ABX_E_INVALID, "No fission neutrons were produced.");
      tallies->calc_gen_values();
      // normalize_weights (src/power_iterator.cpp:538-586)
      double W_pos = 0., W_neg = 0.;
      for (auto& b : next_gen) {
        if (b.wgt > 0.)
          W_pos += b.wgt;
        else
          W_neg -= b.wgt;
      }
      const double w_per_part = (double)N / (W_pos - W_neg);
      for (auto& b : next_gen) b.wgt *= w_per_part;
      if (abx::settings::converged) tallies->record_generation();
      tallies->clear_generation();
      bank.clear();
      for (auto& b : next_gen) {
        bank.push_back(abx::Particle(b.r, b.u, b.E, b.wgt, histories_counter++));
        bank.back().initialize_rng(d.settings.seed, d.settings.stride);
        bank.back().set_family_id(b.family_id);
      }
      if (g == nignored) abx::settings::converged = true;
    }
    std::printf("kcol %.6f %.6f %d\n", tallies->kcol_avg(), tallies->kcol_err(),
                tallies->generations());
    std::printf("kabs %.6f %.6f\n", tallies->kabs_avg(), tallies->kabs_err());
    std::printf("leakage %.6f %.6f\n", tallies->leakage_avg(), tallies->leakage_err());
    std::printf("gens %zu\n", tallies->k_col_vec.size());
  } catch (const abx::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
