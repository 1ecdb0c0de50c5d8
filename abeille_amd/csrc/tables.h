// tables.h - host-side construction of the flattened problem tables.
//
// Validates an abx_problem_desc and derives every table the kernels read,
// with the same arithmetic the reference performs while building its
// objects:
//  * MGNuclide ctor: make_scatter_xs / normalize_chi / check_xs /
//    check_fissile (src/mg_nuclide.cpp:37-333);
//  * the cumulative tables std::discrete_distribution would build per call
//    (built once here, see rng.h);
//  * majorant (src/majorant.cpp:133-176) and carter sampling cross sections
//    (src/carter_tracker.cpp:60-75);
//  * Cell::simplify / check_for_bc (src/cell.cpp:196-232), lattice corner
//    shift (src/rect_lattice.cpp:33-52), has_boundary_conditions
//    (src/cell_universe.cpp:30-41, src/lattice.cpp:35-52);
//  * MeshTally geometry (src/mesh_tally.cpp:62-116) with net weight =
//    nparticles (Tallies::add_*_mesh_tally, src/tallies.cpp:77-87).
#pragma once

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/abx.h"
#include "devtypes.h"
#include "rng.h"

struct HostTables {
  std::vector<DSurf> surf;
  std::vector<DCell> cell;
  std::vector<int32_t> rpn;
  std::vector<DUniv> univ;
  std::vector<int32_t> ucells, elems;
  int32_t root = 0;
  int32_t nmat = 0, G = 0, n_dgrp_max = 1;
  std::vector<double> ebounds, emid, majorant, sampling;
  std::vector<double> Et, Eabs, Ef, nu_tot, nu_d, Eel, vEf;
  std::vector<int32_t> fissile;
  std::vector<double> scat_cdf, chi_cdf;
  std::vector<int32_t> ang_base;  // [nmat] start in ang_off, -1 = isotropic everywhere
  std::vector<int32_t> ang_off;   // per anisotropic material: [G][G] + 1 into ang_mu/pdf/cdf
  std::vector<double> ang_mu, ang_pdf, ang_cdf;
  std::vector<int32_t> dgrp_n;
  std::vector<double> dgrp_cdf;
  std::vector<DTally> tally;
  std::vector<double> tally_eb;
  std::vector<int32_t> tally_gbin;
  int64_t tally_bins = 0;
  std::vector<abx_source> sources;
  int32_t max_pads = 1, max_levels = 1;  // cursor depth bound over all descents
  bool extended = false;  // hexagonal lattices, off-axis cylinders, exact cancellation
  abx_settings st{};
  std::vector<double> chi_norm;  // [nmat][G][G] fission spectra, rows normalised
  // exact cancellation (ExactMGCancelator / BasicExactMGCancelator): the
  // configuration, the regions' pitch and shape (energy regions last, 1
  // without group bins), the group bins as offsets + groups and the bin of
  // each group (the number of bins for a group in none)
  int32_t exact = 0;  // abx_exact_kind, 0 = none
  abx_exact_cancel ex{};
  double ex_pitch[3] = {0., 0., 0.};
  int32_t ex_shape[4] = {1, 1, 1, 1};
  std::vector<int32_t> ex_gb_off, ex_gb, ex_g2bin;
  double min_energy = 0., max_energy = 100000.;
};

struct TableError : std::runtime_error {
  int code;
  TableError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

static inline void tables_require(bool ok, const std::string& m) {
  if (!ok) throw TableError(ABX_E_INVALID, m);
}

// make_cancelator (src/cancelator.cpp:32-78) and the cancelators'
// constructors (src/exact_mg_cancelator.cpp:41-118,
// src/basic_exact_mg_cancelator.cpp:43-66): validation, pitch, group bins.
static inline void build_exact(const abx_problem_desc* d, HostTables& T) {
  const abx_exact_cancel& e = *d->exact_cancel;
  tables_require(e.kind == ABX_EXACT_MG || e.kind == ABX_BASIC_EXACT_MG, "Unkown cancelator type.");
  tables_require(d->cancel == nullptr, "one cancelator per problem");
  tables_require(T.st.tracking != ABX_SURFACE_TRACKING,
                 "exect cancelators may not be used with surface-tracking.");
  T.exact = e.kind;
  T.ex = e;
  tables_require(e.low[0] < e.high[0] && e.low[1] < e.high[1] && e.low[2] < e.high[2],
                 "r_low >= r_hi");
  tables_require(e.n_samples > 0, "n-samples must be greater than zero.");
  for (int k = 0; k < 3; k++) {
    tables_require(e.shape[k] >= 1, "cancelator shape");
    T.ex_shape[k] = e.shape[k];
    T.ex_pitch[k] = (e.high[k] - e.low[k]) / (double)e.shape[k];
  }
  const int G = T.G;
  T.ex_gb_off.assign(1, 0);
  T.ex_gb.clear();
  if (e.kind == ABX_EXACT_MG) {
    tables_require(e.n_group_bins >= 0, "group bins");
    tables_require(e.n_group_bins == 0 || (e.group_bin_offsets && e.group_bins), "group bins");
    std::vector<int> seen((size_t)G, 0);
    T.ex_g2bin.assign((size_t)G, e.n_group_bins);
    for (int b = 0; b < e.n_group_bins; b++) {
      tables_require(e.group_bin_offsets[b] <= e.group_bin_offsets[b + 1], "group bins");
      for (int q = e.group_bin_offsets[b]; q < e.group_bin_offsets[b + 1]; q++) {
        const int g = e.group_bins[q];
        tables_require(g >= 0 && g < G, "grp >= settings::ngroups");
        tables_require(seen[(size_t)g]++ == 0, "Group is included more than once in group_bins.");
        T.ex_gb.push_back(g);
        T.ex_g2bin[(size_t)g] = b;
      }
      tables_require(e.group_bin_offsets[b + 1] > e.group_bin_offsets[b] || !e.chi_matrix,
                     "empty group bin");
      T.ex_gb_off.push_back((int32_t)T.ex_gb.size());
    }
    tables_require(!e.chi_matrix || e.n_group_bins > 0,
                   "Chi matrix is used, but no group_bins provided.");
    T.ex_shape[3] = e.n_group_bins > 0 ? e.n_group_bins : 1;
  } else {
    tables_require(e.beta >= ABX_BETA_ZERO && e.beta <= ABX_BETA_AVERAGE_G,
                   "Unkown beta entry for basic exact MG cancelator.");
    T.ex_g2bin.assign((size_t)G, 0);
    T.ex_shape[3] = 1;
  }
  // region numbers (the mesh widened by one face each way, times the
  // materials) and site positions share one 64-bit sort key, 32 bits each
  const double regions = (double)(T.ex_shape[0] + 1) * (double)(T.ex_shape[1] + 1) *
                         (double)(T.ex_shape[2] + 1) * (double)(T.ex_shape[3] + 1) *
                         (double)std::max(1, T.nmat);
  tables_require(regions < 4294967295.0, "exact cancellation mesh too large");
  T.extended = true;  // the parent records come from the extended kernels
}

static inline void build_tables(const abx_problem_desc* d, HostTables& T) {
  tables_require(d != nullptr, "null description");
  const abx_settings& s = d->settings;
  T.st = s;
  const int G = s.ngroups;
  tables_require(G >= 1 && s.energy_bounds != nullptr, "ngroups / energy-bounds");
  tables_require(s.nparticles > 0, "nparticles");
  T.G = G;
  T.ebounds.assign(s.energy_bounds, s.energy_bounds + G + 1);
  for (int g = 0; g < G; g++) tables_require(T.ebounds[g] < T.ebounds[g + 1], "energy-bounds not increasing");
  T.emid.resize(G);
  for (int g = 0; g < G; g++) T.emid[g] = 0.5 * (T.ebounds[g] + T.ebounds[g + 1]);

  // surfaces
  tables_require(d->n_surfaces >= 1, "no surfaces");
  std::vector<std::pair<int, DSurf>> aux;  // companions of off-axis cylinders
  for (int i = 0; i < d->n_surfaces; i++) {
    const abx_surface& a = d->surfaces[i];
    tables_require(a.type >= 0 && a.type <= 8, "surface type");
    tables_require(a.boundary >= 0 && a.boundary <= 2, "surface boundary");
    DSurf S{};
    for (int k = 0; k < 4; k++) S.p[k] = a.p[k];
    S.type = a.type;
    if (a.type == ABX_CYL) {
      // Cylinder (src/cylinder.cpp:30-117) is the quadric
      // alpha x^2 + beta y^2 + gamma z^2 = R^2 with alpha = 1 - u0^2 etc. of
      // the normalised axis.  Along a coordinate axis two of alpha, beta,
      // gamma are exactly 1 and the third exactly 0, and its sign, distance
      // and normal evaluate bit for bit as the x/y/z-cylinder's (the zero
      // terms add nothing, the others add in either order), so it is
      // stored as one.  Off the axes (the reference's quadric is then not a
      // circular cylinder, but it is what the reference evaluates) it stays
      // ABX_CYL with p = {x0, y0, z0, R} and alpha, beta, gamma in a
      // companion entry appended after the deck's surfaces; only the
      // extended-geometry kernels evaluate it.
      double u0 = a.p[3], v0 = a.p[4], w0 = a.p[5];
      const double mag = sqrt(u0 * u0 + v0 * v0 + w0 * w0);
      tables_require(mag != 0., "Direction vector of Cylinder can not have a mangitude of 0.");
      u0 /= mag;
      v0 /= mag;
      w0 /= mag;
      const double R = a.p[6];
      if (v0 == 0. && w0 == 0.) {
        S.type = ABX_XCYL;
        S.p[0] = a.p[1], S.p[1] = a.p[2], S.p[2] = R, S.p[3] = 0.;
      } else if (u0 == 0. && w0 == 0.) {
        S.type = ABX_YCYL;
        S.p[0] = a.p[0], S.p[1] = a.p[2], S.p[2] = R, S.p[3] = 0.;
      } else if (u0 == 0. && v0 == 0.) {
        S.type = ABX_ZCYL;
        S.p[0] = a.p[0], S.p[1] = a.p[1], S.p[2] = R, S.p[3] = 0.;
      } else {
        S.p[0] = a.p[0], S.p[1] = a.p[1], S.p[2] = a.p[2], S.p[3] = R;
        DSurf Q{};
        Q.type = ABX_CYL;
        Q.bc = ABX_NORMAL;
        Q.p[0] = 1. - u0 * u0;
        Q.p[1] = 1. - v0 * v0;
        Q.p[2] = 1. - w0 * w0;
        aux.push_back({i, Q});
        T.extended = true;
      }
    }
    S.bc = a.boundary;
    T.surf.push_back(S);
  }
  for (const auto& q : aux) {
    T.surf[q.first].pad_[0] = (int32_t)T.surf.size() - q.first;
    T.surf.push_back(q.second);
  }
  // cells
  for (int i = 0; i < d->n_cells; i++) {
    const abx_cell& a = d->cells[i];
    tables_require(a.rpn_offset >= 0 && a.rpn_len >= 1 && a.rpn_offset + a.rpn_len <= d->n_rpn, "cell rpn range");
    std::vector<int32_t> rpn(d->rpn + a.rpn_offset, d->rpn + a.rpn_offset + a.rpn_len);
    bool simple = true;
    for (int32_t el : rpn)
      if (el == ABX_OP_COMP || el == ABX_OP_UNIN) simple = false;
    if (simple) {
      std::vector<int32_t> t;
      for (int32_t el : rpn)
        if (el < ABX_OP_UNIN) t.push_back(el);
      rpn = t;
    }
    int depth = 0, maxdepth = 0;
    bool bc = false;
    for (int32_t tok : rpn) {
      if (tok >= ABX_OP_UNIN) {
        if (tok != ABX_OP_COMP) depth--;
        continue;
      }
      tables_require(tok != 0 && std::abs(tok) <= d->n_surfaces, "cell token out of range");
      depth++;
      maxdepth = std::max(maxdepth, depth);
      const int b = d->surfaces[std::abs(tok) - 1].boundary;
      if (b == ABX_VACUUM || b == ABX_REFLECTIVE) bc = true;
    }
    tables_require(maxdepth <= 64, "cell region too deep");
    DCell C{};
    C.rpn_off = (int32_t)T.rpn.size();
    C.rpn_len = (int32_t)rpn.size();
    C.simple = simple ? 1 : 0;
    C.fill_type = a.fill_type;
    C.fill = a.fill;
    C.has_bc = bc ? 1 : 0;
    C.id = a.id;
    if (a.fill_type == ABX_FILL_MATERIAL)
      tables_require(a.fill >= 0 && a.fill < d->n_materials, "cell material index");
    else
      tables_require(a.fill >= 0 && a.fill < d->n_universes, "cell universe index");
    T.rpn.insert(T.rpn.end(), rpn.begin(), rpn.end());
    T.cell.push_back(C);
  }
  // re-descent universe of a cell pad: the universe whose id equals the
  // cell's id, else index 0 (Tracker::get_current, tracker.hpp:285-286)
  for (auto& C : T.cell) {
    C.redescend = 0;
    for (int u = 0; u < d->n_universes; u++)
      if (d->universes[u].id == C.id) {
        C.redescend = u;
        break;
      }
  }
  // universes
  tables_require(d->n_universes >= 1, "no universes");
  for (int i = 0; i < d->n_universes; i++) {
    const abx_universe& a = d->universes[i];
    DUniv U{};
    U.type = a.type;
    U.outer = -1;
    if (a.type == ABX_CELL_UNIVERSE) {
      tables_require(a.n_cells >= 1 && a.cells_offset >= 0 && a.cells_offset + a.n_cells <= d->n_universe_cells, "universe cells");
      U.cells_off = (int32_t)T.ucells.size();
      U.ncells = a.n_cells;
      for (int k = 0; k < a.n_cells; k++) {
        const int ci = d->universe_cells[a.cells_offset + k];
        tables_require(ci >= 0 && ci < d->n_cells, "universe cell index");
        T.ucells.push_back(ci);
      }
    } else if (a.type == ABX_HEX_LATTICE) {
      // HexLattice constructor and set_elements (src/hex_lattice.cpp:30-56,
      // 219-243); make_hex_lattice's checks (src/hex_lattice.cpp:469-503)
      tables_require(a.shape[0] >= 1 && a.shape[1] >= 1, "hex lattice shape");
      tables_require(a.pitch[0] > 0. && a.pitch[1] > 0., "hex lattice pitch");
      tables_require(a.hex_top == ABX_HEX_POINTY || a.hex_top == ABX_HEX_FLAT, "hex lattice top");
      tables_require(a.shape[0] <= 4096 && a.shape[1] <= (1 << 20), "hex lattice size");
      const uint32_t rings = (uint32_t)a.shape[0], width = 2 * (rings - 1) + 1;
      const uint32_t mid = width / 2;
      T.extended = true;
      U.rings = (int32_t)rings;
      U.top = a.hex_top;
      U.Nx = U.Ny = (int32_t)width;
      U.Nz = a.shape[1];
      U.Px = a.pitch[0];
      U.Pz = a.pitch[1];
      U.Xl = a.origin[0];
      U.Yl = a.origin[1];
      U.Zl = a.origin[2];
      U.z_low = U.Zl - 0.5 * (double)(uint32_t)U.Nz * U.Pz;
      U.c6 = std::cos(ABX_PI / 6.0);
      U.s6 = std::sin(ABX_PI / 6.0);
      U.c3 = std::cos(ABX_PI / 3.0);
      U.s3 = std::sin(ABX_PI / 3.0);
      const int64_t nhex = 3 * (int64_t)rings * (rings - 1) + 1;
      const int64_t ne = nhex * U.Nz;
      tables_require(a.elems_offset >= 0 && a.elems_offset + ne <= d->n_lattice_elems,
                     "Improper number of universes for HexLattice.");
      U.elems_off = (int32_t)T.elems.size();
      T.elems.resize(T.elems.size() + (size_t)width * width * U.Nz, -1);
      int64_t k = 0;
      for (uint32_t az = 0; az < (uint32_t)U.Nz; az++)
        for (uint32_t ar = 0; ar < width; ar++)
          for (uint32_t aq = 0; aq < width; aq++) {
            const int32_t q = (int32_t)(aq - mid), r = (int32_t)(ar - mid), y = -q - r;
            const uint32_t ring = std::max(std::max((uint32_t)std::abs(q), (uint32_t)std::abs(y)),
                                           (uint32_t)std::abs(r));
            if (ring >= rings) continue;
            const int e = d->lattice_elems[a.elems_offset + k++];
            tables_require(e >= -1 && e < d->n_universes, "lattice element index");
            T.elems[U.elems_off + ((size_t)az * width + ar) * width + aq] = e;
          }
      tables_require(a.outer >= -1 && a.outer < d->n_universes, "lattice outer");
      U.outer = a.outer;
    } else {
      tables_require(a.type == ABX_RECT_LATTICE, "universe type");
      for (int k = 0; k < 3; k++) tables_require(a.shape[k] >= 1 && a.pitch[k] > 0., "lattice shape/pitch");
      U.Nx = a.shape[0];
      U.Ny = a.shape[1];
      U.Nz = a.shape[2];
      U.Px = a.pitch[0];
      U.Py = a.pitch[1];
      U.Pz = a.pitch[2];
      U.Pxi = 1. / U.Px;
      U.Pyi = 1. / U.Py;
      U.Pzi = 1. / U.Pz;
      U.Xl = a.origin[0] - (double)(uint32_t)U.Nx * 0.5 * U.Px;
      U.Yl = a.origin[1] - (double)(uint32_t)U.Ny * 0.5 * U.Py;
      U.Zl = a.origin[2] - (double)(uint32_t)U.Nz * 0.5 * U.Pz;
      const int64_t ne = (int64_t)U.Nx * U.Ny * U.Nz;
      tables_require(a.elems_offset >= 0 && a.elems_offset + ne <= d->n_lattice_elems, "lattice elements");
      U.elems_off = (int32_t)T.elems.size();
      for (int64_t k = 0; k < ne; k++) {
        const int e = d->lattice_elems[a.elems_offset + k];
        tables_require(e >= -1 && e < d->n_universes, "lattice element index");
        T.elems.push_back(e);
      }
      tables_require(a.outer >= -1 && a.outer < d->n_universes, "lattice outer");
      U.outer = a.outer;
    }
    T.univ.push_back(U);
  }
  for (auto& U : T.univ)
    if (U.type == ABX_CELL_UNIVERSE)
      for (int k = 0; k < U.ncells; k++)
        if (T.cell[T.ucells[U.cells_off + k]].has_bc) U.has_bc = 1;
  for (int pass = 0; pass < 8; pass++)
    for (auto& U : T.univ)
      if (U.type != ABX_CELL_UNIVERSE && U.outer >= 0) U.has_bc = T.univ[U.outer].has_bc;
  tables_require(d->root_universe >= 0 && d->root_universe < d->n_universes, "root universe");
  T.root = d->root_universe;
  // cursor depth: pads along the deepest descent and coordinate frames
  // (a lattice element opens a new frame).  Cycles are rejected like the
  // reference's contains_universe check (src/parser.cpp:206-214).
  {
    const int NU = (int)T.univ.size();
    std::vector<int> pads(NU, -1), lvls(NU, -1), state(NU, 0);
    std::function<void(int)> visit = [&](int ui) {
      if (state[ui] == 2) return;
      tables_require(state[ui] == 0, "Cyclical universe reference found");
      state[ui] = 1;
      const DUniv& U = T.univ[ui];
      int pd = 0, lv = 1;
      if (U.type == ABX_CELL_UNIVERSE) {
        pd = 2;
        for (int k = 0; k < U.ncells; k++) {
          const DCell& c = T.cell[T.ucells[U.cells_off + k]];
          if (c.fill_type == ABX_FILL_UNIVERSE) {
            visit(c.fill);
            pd = std::max(pd, 2 + pads[c.fill]);
            lv = std::max(lv, lvls[c.fill]);
          }
        }
      } else {
        pd = 1;
        const int64_t ne = (int64_t)U.Nx * U.Ny * U.Nz;
        for (int64_t k = 0; k < ne; k++) {
          const int e = T.elems[U.elems_off + k];
          if (e >= 0) {
            visit(e);
            pd = std::max(pd, 1 + pads[e]);
            lv = std::max(lv, 1 + lvls[e]);
          }
        }
        if (U.outer >= 0) {
          visit(U.outer);
          pd = std::max(pd, 1 + pads[U.outer]);
          lv = std::max(lv, lvls[U.outer]);
        }
      }
      pads[ui] = pd;
      lvls[ui] = lv;
      state[ui] = 2;
    };
    visit(T.root);
    T.max_pads = pads[T.root];
    T.max_levels = lvls[T.root];
    tables_require(T.max_pads <= ABX_MAXPADS, "geometry nests deeper than the cursor supports");
    tables_require(T.max_levels <= ABX_MAXLVL, "too many nested lattices");
  }

  // materials
  const int M = d->n_materials;
  tables_require(M >= 1, "no materials");
  T.nmat = M;
  const size_t MG = (size_t)M * G, MGG = MG * G;
  T.Et.assign(MG, 0.);
  T.Eabs.assign(MG, 0.);
  T.Ef.assign(MG, 0.);
  T.nu_tot.assign(MG, 0.);
  T.nu_d.assign(MG, 0.);
  T.Eel.assign(MG, 0.);
  T.vEf.assign(MG, 0.);
  T.fissile.assign(M, 0);
  T.scat_cdf.assign(MGG, 0.);
  T.chi_cdf.assign(MGG, 0.);
  T.ang_base.assign(M, -1);
  T.ang_off.clear();
  T.dgrp_n.assign(M, 0);
  int ndmax = 1;
  for (int m = 0; m < M; m++) ndmax = std::max(ndmax, d->materials[m].n_delayed);
  T.n_dgrp_max = ndmax;
  T.dgrp_cdf.assign((size_t)M * ndmax, 1.0);
  for (int m = 0; m < M; m++) {
    const abx_material& a = d->materials[m];
    tables_require(a.total && a.absorption && a.scatter, "material arrays");
    std::vector<double> Ef(G, 0.), nup(G, 0.), nud(G, 0.);
    bool fis = false;
    if (a.fission) {
      for (int g = 0; g < G; g++) {
        Ef[g] = a.fission[g];
        tables_require(Ef[g] >= 0., "Negative fission xs");
        if (Ef[g] > 0.) fis = true;
      }
    }
    if (fis) {
      tables_require(a.nu_prompt != nullptr, "No nu data is provided");
      for (int g = 0; g < G; g++) nup[g] = a.nu_prompt[g];
      if (a.nu_delayed)
        for (int g = 0; g < G; g++) nud[g] = a.nu_delayed[g];
    }
    // chi, normalised per row (normalize_chi)
    std::vector<double> chi((size_t)G * G, 0.);
    if (fis) {
      tables_require(a.chi != nullptr, "Invalid chi matrix entry");
      for (int k = 0; k < G * G; k++) chi[k] = a.chi[k];
    }
    for (int i = 0; i < G; i++) {
      double chi_i = 0.;
      for (int o = 0; o < G; o++) chi_i += chi[(size_t)i * G + o];
      for (int o = 0; o < G; o++) chi[(size_t)i * G + o] /= chi_i;
    }
    T.chi_norm.insert(T.chi_norm.end(), chi.begin(), chi.end());
    // scatter: Es row sums, Ps rows normalised (make_scatter_xs)
    std::vector<double> Ps((size_t)G * G), Es(G);
    for (int k = 0; k < G * G; k++) {
      Ps[k] = a.scatter[k];
      tables_require(Ps[k] >= 0., "Negative scattering component");
    }
    for (int i = 0; i < G; i++) {
      Es[i] = 0.;
      for (int o = 0; o < G; o++) Es[i] += Ps[(size_t)i * G + o];
      for (int o = 0; o < G; o++) Ps[(size_t)i * G + o] /= Es[i];
    }
    for (int g = 0; g < G; g++) {
      const size_t k = (size_t)m * G + g;
      const double Ea = a.absorption[g];
      tables_require(Ea >= 0., "Ea negative");
      tables_require(Ef[g] <= Ea, "Ef > Ea");
      const double diff = a.total[g] - (Es[g] + Ea);
      tables_require(fabs(diff) / a.total[g] <= 0.001, "Es + Ea != Et");
      T.Et[k] = a.total[g];
      T.Ef[k] = Ef[g];
      T.Eabs[k] = Ef[g] + (Ea - Ef[g]);
      T.nu_tot[k] = nup[g] + nud[g];
      T.nu_d[k] = nud[g];
      T.Eel[k] = Es[g];
      T.vEf[k] = 1.0 * T.nu_tot[k] * Ef[g];
      build_discrete_cdf(&Ps[(size_t)g * G], G, &T.scat_cdf[k * G]);
      build_discrete_cdf(&chi[(size_t)g * G], G, &T.chi_cdf[k * G]);
      if (Ef[g] != 0. && (nup[g] + nud[g]) > 0.) T.fissile[m] = 1;
    }
    // angle tables.  A material without tables is isotropic everywhere
    // (MGAngleDistribution(), src/mg_angle_distribution.cpp:31-32): its
    // cosine is sampled in closed form, so it stores nothing.
    if (a.angle_offsets) {
      T.ang_base[m] = (int32_t)T.ang_off.size();
      for (int gi = 0; gi < G; gi++)
        for (int go = 0; go < G; go++) {
          T.ang_off.push_back((int32_t)T.ang_mu.size());
          const int b0 = a.angle_offsets[gi * G + go], b1 = a.angle_offsets[gi * G + go + 1];
          tables_require(b1 - b0 >= 2, "angle table too short");
          const size_t first = T.ang_mu.size();
          for (int q = b0; q < b1; q++) {
            T.ang_mu.push_back(a.angle_mu[q]);
            T.ang_pdf.push_back(a.angle_pdf[q]);
            T.ang_cdf.push_back(a.angle_cdf[q]);
          }
          tables_require(T.ang_cdf[first] == 0. && T.ang_cdf.back() == 1., "angle cdf must run 0..1");
        }
      T.ang_off.push_back((int32_t)T.ang_mu.size());
    }
    T.dgrp_n[m] = a.n_delayed;
    if (a.n_delayed > 0) {
      tables_require(a.delayed_probabilities != nullptr, "delayed probabilities");
      build_discrete_cdf(a.delayed_probabilities, a.n_delayed, &T.dgrp_cdf[(size_t)m * ndmax]);
    }
  }
  // energy limits (src/parser.cpp:152-167)
  T.min_energy = std::max(0., T.ebounds.front());
  T.max_energy = std::min(100000., T.ebounds.back());
  // majorant and carter sampling xs
  T.majorant.assign(G, 0.);
  for (int m = 0; m < M; m++)
    for (int g = 0; g < G; g++) {
      const double xs = T.Et[(size_t)m * G + g];
      if (xs > T.majorant[g]) T.majorant[g] = xs;
    }
  T.sampling = T.majorant;
  if (s.tracking == ABX_CARTER_TRACKING) {
    tables_require(s.sample_xs_ratio != nullptr, "carter tracking needs sample_xs_ratio");
    for (int g = 0; g < G; g++) T.sampling[g] = T.majorant[g] * s.sample_xs_ratio[g];
  }
  // sources
  tables_require(d->n_sources >= 1, "no sources");
  for (int i = 0; i < d->n_sources; i++) {
    const abx_source& a = d->sources[i];
    tables_require(a.spatial == ABX_SRC_POINT || a.spatial == ABX_SRC_BOX, "source spatial");
    tables_require(a.direction == ABX_DIR_ISOTROPIC || a.direction == ABX_DIR_MONO, "source direction");
    T.sources.push_back(a);
  }
  // tallies
  for (int i = 0; i < d->n_tallies; i++) {
    const abx_mesh_tally& a = d->tallies[i];
    tables_require(a.estimator == ABX_EST_COLLISION || a.estimator == ABX_EST_TRACK_LENGTH, "estimator");
    tables_require(a.shape[0] >= 1 && a.shape[1] >= 1 && a.shape[2] >= 1 && a.n_ebounds >= 2, "tally shape");
    DTally D{};
    D.estimator = a.estimator;
    D.quantity = a.quantity;
    D.Nx = a.shape[0];
    D.Ny = a.shape[1];
    D.Nz = a.shape[2];
    D.Ne = a.n_ebounds - 1;
    D.eb_off = (int32_t)T.tally_eb.size();
    T.tally_eb.insert(T.tally_eb.end(), a.ebounds, a.ebounds + a.n_ebounds);
    D.data_off = T.tally_bins;
    T.tally_bins += (int64_t)D.Ne * D.Nx * D.Ny * D.Nz;
    D.lx = a.low[0];
    D.ly = a.low[1];
    D.lz = a.low[2];
    D.hx = a.high[0];
    D.hy = a.high[1];
    D.hz = a.high[2];
    D.dx = (D.hx - D.lx) / (double)D.Nx;
    D.dy = (D.hy - D.ly) / (double)D.Ny;
    D.dz = (D.hz - D.lz) / (double)D.Nz;
    D.dxi = 1. / D.dx;
    D.dyi = 1. / D.dy;
    D.dzi = 1. / D.dz;
    D.net_weight = (double)s.nparticles;
    T.tally.push_back(D);
    for (int g = 0; g < G; g++) {
      int l = -1;
      const double E = T.emid[g];
      for (int e = 0; e < D.Ne; e++)
        if (a.ebounds[e] <= E && E <= a.ebounds[e + 1]) {
          l = e;
          break;
        }
      T.tally_gbin.push_back(l);
    }
  }
  if (T.tally_gbin.empty()) T.tally_gbin.push_back(-1);
  if (T.tally_eb.empty()) T.tally_eb.push_back(0.);
  if (d->exact_cancel) build_exact(d, T);
}

// DevProblem view of host tables (pointers into the vectors, or into device
// copies when `dev` maps each host array to its device address).
template <class Map>
static inline DevProblem dev_view(const HostTables& T, Map map, double* tally_data) {
  DevProblem P{};
  P.surf = map(T.surf);
  P.cell = map(T.cell);
  P.rpn = map(T.rpn);
  P.univ = map(T.univ);
  P.ucells = map(T.ucells);
  P.elems = map(T.elems);
  P.root = T.root;
  P.nmat = T.nmat;
  P.G = T.G;
  P.mode = T.st.mode;
  P.tracking = T.st.tracking;
  P.ntally = (int32_t)T.tally.size();
  P.n_dgrp_max = T.n_dgrp_max;
  P.precision = T.st.precision;
  P.trk_D = T.max_pads;
  P.trk_L = T.max_levels;
  P.wgt_cutoff = T.st.wgt_cutoff;
  P.wgt_survival = T.st.wgt_survival;
  P.wgt_split = T.st.wgt_split;
  P.min_energy = T.min_energy;
  P.max_energy = T.max_energy;
  P.seed = T.st.seed;
  P.stride = T.st.stride;
  P.ebounds = map(T.ebounds);
  P.emid = map(T.emid);
  P.majorant = map(T.majorant);
  P.sampling = map(T.sampling);
  P.Et = map(T.Et);
  P.Eabs = map(T.Eabs);
  P.Ef = map(T.Ef);
  P.nu_tot = map(T.nu_tot);
  P.nu_d = map(T.nu_d);
  P.Eel = map(T.Eel);
  P.vEf = map(T.vEf);
  P.fissile = map(T.fissile);
  P.scat_cdf = map(T.scat_cdf);
  P.chi_cdf = map(T.chi_cdf);
  P.ang_base = map(T.ang_base);
  P.ang_off = map(T.ang_off);
  P.ang_mu = map(T.ang_mu);
  P.ang_pdf = map(T.ang_pdf);
  P.ang_cdf = map(T.ang_cdf);
  P.dgrp_n = map(T.dgrp_n);
  P.dgrp_cdf = map(T.dgrp_cdf);
  P.tally = map(T.tally);
  P.tally_eb = map(T.tally_eb);
  P.tally_data = tally_data;
  P.tally_gbin = map(T.tally_gbin);
  return P;
}
