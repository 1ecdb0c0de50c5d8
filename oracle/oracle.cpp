// oracle.cpp - CPU restatement of Abeille's multigroup transport path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): the checker the HIP path is
// compared with, and the CPU baseline.  Written from the reference sources;
// every function names the reference lines it restates.  Structure follows
// the reference on purpose (history-based OpenMP loop, a Tracker with a
// vector of lily pads, a discrete_distribution built per sample) so that it
// behaves - and costs - like the reference's CPU path.
//
// Parity status of each part (how it is pinned, see tests/):
//  * pcg32 / rand / exponential / discrete: golden vectors from pcg-cpp and
//    libstdc++ (tests/golden/rng_golden.json) + O'Neill's demo output.
//  * log/sin/cos: glibc, or the portable detmath.h checked <= 1 ulp to glibc.
//  * surfaces, cells, lattices, tracker: closed-form distances, brute-force
//    point location, and face tie-break cases (tests/test_oracle_geometry.py);
//    hexagonal lattices: brute-force nearest-centre location, tile distances
//    against ray/line intersections and closed forms, k-infinity through
//    the lattice (tests/test_hex_lattice.py).
//  * collision physics: closed forms (infinite-medium fluxes for delta,
//    surface and carter tracking - carter on both of its branches -,
//    k-infinity, leakage + absorption balance) and the LA-13511 analytic k
//    values (tests/test_oracle_physics.py, tests/test_carter_closed_form.py).
//  * mesh tallies: track-length scores of fixed rays through asymmetric
//    meshes against per-cell lengths from sorted plane crossings, and
//    collision scores against the sites in the event traces, bin by bin
//    (tests/test_tally_pinning.py).
//  * generation statistics: running means and errors against numpy over the
//    same per-generation values (tests/test_statistics_recurrences.py).
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <random>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../abeille_amd/csrc/detmath.h"

namespace orc {

// include/utils/constants.hpp:42-57
constexpr double INF = DBL_MAX;
constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double SURFACE_COINCIDENT = 1E-12;
constexpr double BOUNDRY_TOL = 500. * SURFACE_COINCIDENT;

static thread_local std::string g_err;
static bool g_portable = true;

static double SIN(double x) { return g_portable ? dm_sin(x) : std::sin(x); }
static double COS(double x) { return g_portable ? dm_cos(x) : std::cos(x); }
static double EXP(double x) { return g_portable ? dm_exp(x) : std::exp(x); }

struct Fatal {
  int code;
  std::string msg;
};

[[noreturn]] static void fatal(int code, const std::string& m) {
  throw Fatal{code, m};
}

//==========================================================================
// Vector math: include/utils/{vector,position,direction}.hpp
struct V3 {
  double x, y, z;
};
static inline V3 add(const V3& a, const V3& b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
static inline V3 sub(const V3& a, const V3& b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
static inline V3 scale(const V3& v, double d) {
  return {v.x * d, v.y * d, v.z * d};
}
static inline double dot(const V3& a, const V3& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
static inline double norm(const V3& a) {
  return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
// Direction(x, y, z): normalising constructor (direction.hpp:37-42)
static inline V3 direction(double x, double y, double z) {
  V3 v{x, y, z};
  double m = norm(v);
  return {x / m, y / m, z / m};
}
// Direction(mu, phi) (direction.hpp:43-62)
static inline V3 direction_mu_phi(double mu, double phi) {
  if (mu < -1.)
    mu = -1.;
  else if (mu > 1.)
    mu = 1.;
  if (phi < 0.)
    phi = 0.;
  else if (phi > 2 * PI)
    phi = 2 * PI;
  double x = std::sqrt(1. - mu * mu) * COS(phi);
  double y = std::sqrt(1. - mu * mu) * SIN(phi);
  double z = mu;
  return direction(x, y, z);
}
// rotate_direction (direction.hpp:130-151)
static inline V3 rotate_direction(V3 u, double mu, double phi) {
  double c = COS(phi);
  double s = SIN(phi);
  double sqrt_mu = std::sqrt(1. - mu * mu);
  double sqrt_w = std::sqrt(1. - u.z * u.z);
  double ux, uy, uz;
  if (sqrt_w > 1.E-10) {
    ux = mu * u.x + sqrt_mu * (u.x * u.z * c - u.y * s) / sqrt_w;
    uy = mu * u.y + sqrt_mu * (u.y * u.z * c + u.x * s) / sqrt_w;
    uz = mu * u.z - sqrt_mu * sqrt_w * c;
  } else {
    double sqrt_v = std::sqrt(1. - u.y * u.y);
    ux = mu * u.x + sqrt_mu * (u.x * u.y * c + u.z * s) / sqrt_v;
    uy = mu * u.y - sqrt_mu * sqrt_v * c;
    uz = mu * u.z + sqrt_mu * (u.y * u.z * c - u.x * s) / sqrt_v;
  }
  return direction(ux, uy, uz);
}

//==========================================================================
// pcg32 restated from pcg-cpp (pcg_random.hpp: engine ctor, bump, XSH-RR
// output of the previous state, Brown advance), usable as a
// UniformRandomBitGenerator for libstdc++'s distributions.
struct Pcg32 {
  using result_type = uint32_t;
  static constexpr uint64_t MULT = 6364136223846793005ULL;
  static constexpr uint64_t INC = 1442695040888963407ULL;
  uint64_t state = 0;
  uint64_t calls = 0;  // 32-bit outputs drawn
  static constexpr result_type min() { return 0u; }
  static constexpr result_type max() { return 0xffffffffu; }
  void seed(uint64_t s) { state = (s + INC) * MULT + INC; }
  void advance(uint64_t delta) {
    uint64_t cm = MULT, cp = INC, am = 1, ap = 0;
    while (delta > 0) {
      if (delta & 1) {
        am *= cm;
        ap = ap * cm + cp;
      }
      cp = (cm + 1) * cp;
      cm *= cm;
      delta >>= 1;
    }
    state = am * state + ap;
  }
  result_type operator()() {
    uint64_t old = state;
    state = old * MULT + INC;
    calls++;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  }
};

// RNG (include/utils/rng.hpp:41-96), on libstdc++'s own distributions.
struct RNG {
  static double rand(Pcg32& g) {
    std::uniform_real_distribution<double> unit(0., 1.);
    return unit(g);
  }
  static double exponential(Pcg32& g, double lambda) {
    if (lambda == 0.) return INF;
    if (!g_portable) {
      std::exponential_distribution<double> d(lambda);
      return d(g);
    }
    // same draw as exponential_distribution, with the portable log
    double u = std::generate_canonical<double, 53>(g);
    return -dm_log(1.0 - u) / lambda;
  }
  static int discrete(Pcg32& g, const double* b, const double* e) {
    std::discrete_distribution<int> d(b, e);
    return d(g);
  }
};

//==========================================================================
// Surfaces (src/{x,y,z}plane.cpp, src/plane.cpp, src/{x,y,z}cylinder.cpp,
// src/sphere.cpp, src/cylinder.cpp)
struct Surface {
  int type;
  int bc;
  uint32_t id;
  double p[4];
  // Cylinder (src/cylinder.cpp:30-57): point, normalised-axis coefficients
  double x0 = 0., y0 = 0., z0 = 0., alpha = 0., beta = 0., gamma = 0., R = 0.;

  V3 norm(const V3& r) const {
    switch (type) {
      case ABX_XPLANE:
        return direction(1., 0., 0.);
      case ABX_YPLANE:
        return direction(0., 1., 0.);
      case ABX_ZPLANE:
        return direction(0., 0., 1.);
      case ABX_PLANE:
        return direction(p[0], p[1], p[2]);
      case ABX_XCYL:
        return direction(0., r.y - p[0], r.z - p[1]);
      case ABX_YCYL:
        return direction(r.x - p[0], 0., r.z - p[1]);
      case ABX_ZCYL:
        return direction(r.x - p[0], r.y - p[1], 0.);
      case ABX_CYL:  // Cylinder::norm, src/cylinder.cpp:115-117
        return direction(alpha * (r.x - x0), beta * (r.y - y0), gamma * (r.z - z0));
      default:
        return direction(r.x - p[0], r.y - p[1], r.z - p[2]);
    }
  }

  int sign(const V3& r, const V3& u) const {
    double eval;
    switch (type) {
      case ABX_XPLANE:
        eval = r.x - p[0];
        break;
      case ABX_YPLANE:
        eval = r.y - p[0];
        break;
      case ABX_ZPLANE:
        eval = r.z - p[0];
        break;
      case ABX_PLANE:
        eval = p[0] * r.x + p[1] * r.y + p[2] * r.z - p[3];
        break;
      case ABX_XCYL: {
        const double y = r.y - p[0], z = r.z - p[1];
        eval = y * y + z * z - p[2] * p[2];
        break;
      }
      case ABX_YCYL: {
        const double x = r.x - p[0], z = r.z - p[1];
        eval = x * x + z * z - p[2] * p[2];
        break;
      }
      case ABX_ZCYL: {
        const double x = r.x - p[0], y = r.y - p[1];
        eval = y * y + x * x - p[2] * p[2];
        break;
      }
      case ABX_CYL: {  // Cylinder::sign, src/cylinder.cpp:59-72
        const double x = r.x - x0, y = r.y - y0, z = r.z - z0;
        eval = alpha * x * x + beta * y * y + gamma * z * z - R * R;
        break;
      }
      default: {
        const double x = r.x - p[0], y = r.y - p[1], z = r.z - p[2];
        eval = (x * x) + (y * y) + (z * z) - p[3] * p[3];
        break;
      }
    }
    if (eval > SURFACE_COINCIDENT)
      return 1;
    else if (eval < -SURFACE_COINCIDENT)
      return -1;
    else
      return dot(u, norm(r)) > 0. ? 1 : -1;
  }

  double plane_axis(double r0, double u0, bool on_surf) const {
    const double diff = p[0] - r0;
    if (on_surf || std::abs(diff) < SURFACE_COINCIDENT || u0 == 0.)
      return INF;
    else if (diff / u0 < 0.)
      return INF;
    else
      return diff / u0;
  }

  // a = u1^2 + u2^2 etc. in the reference's operand order per type
  static double cyl(double a, double k, double c, bool on_surf) {
    const double quad = k * k - a * c;
    if (quad < 0.)
      return INF;
    else if (on_surf || std::abs(c) < SURFACE_COINCIDENT) {
      if (k >= 0.)
        return INF;
      else
        return (-k + std::sqrt(quad)) / a;
    } else if (c < 0.) {
      return (-k + std::sqrt(quad)) / a;
    } else {
      const double d = (-k - std::sqrt(quad)) / a;
      if (d < 0.)
        return INF;
      else
        return d;
    }
  }

  double distance(const V3& r, const V3& u, bool on_surf) const {
    switch (type) {
      case ABX_XPLANE:
        return plane_axis(r.x, u.x, on_surf);
      case ABX_YPLANE:
        return plane_axis(r.y, u.y, on_surf);
      case ABX_ZPLANE:
        return plane_axis(r.z, u.z, on_surf);
      case ABX_PLANE: {
        const double num = p[3] - p[0] * r.x - p[1] * r.y - p[2] * r.z;
        const double denom = p[0] * u.x + p[1] * u.y + p[2] * u.z;
        const double d = num / denom;
        if (on_surf || std::abs(d) < SURFACE_COINCIDENT || denom == 0.)
          return INF;
        else if (d < 0.)
          return INF;
        else
          return d;
      }
      case ABX_XCYL: {
        const double a = u.y * u.y + u.z * u.z;
        if (a == 0.) return INF;
        const double y = r.y - p[0], z = r.z - p[1];
        const double k = y * u.y + z * u.z;
        const double c = y * y + z * z - p[2] * p[2];
        return cyl(a, k, c, on_surf);
      }
      case ABX_YCYL: {
        const double a = u.x * u.x + u.z * u.z;
        if (a == 0.) return INF;
        const double x = r.x - p[0], z = r.z - p[1];
        const double k = x * u.x + z * u.z;
        const double c = x * x + z * z - p[2] * p[2];
        return cyl(a, k, c, on_surf);
      }
      case ABX_ZCYL: {
        const double a = u.y * u.y + u.x * u.x;
        if (a == 0.) return INF;
        const double x = r.x - p[0], y = r.y - p[1];
        const double k = y * u.y + x * u.x;
        const double c = y * y + x * x - p[2] * p[2];
        return cyl(a, k, c, on_surf);
      }
      case ABX_CYL: {  // Cylinder::distance, src/cylinder.cpp:74-113
        const double a = alpha * u.x * u.x + beta * u.y * u.y + gamma * u.z * u.z;
        if (a == 0.) return INF;
        const double x = r.x - x0, y = r.y - y0, z = r.z - z0;
        const double k = alpha * x * u.x + beta * y * u.y + gamma * z * u.z;
        const double c = alpha * x * x + beta * y * y + gamma * z * z - R * R;
        return cyl(a, k, c, on_surf);
      }
      default: {  // sphere (src/sphere.cpp:51-76)
        const double x = r.x - p[0], y = r.y - p[1], z = r.z - p[2];
        const double k = x * u.x + y * u.y + z * u.z;
        const double c = x * x + y * y + z * z - p[3] * p[3];
        const double quad = k * k - c;
        if (quad < 0.)
          return INF;
        else if (on_surf || std::abs(c) < SURFACE_COINCIDENT) {
          if (k >= 0.)
            return INF;
          else
            return -k + std::sqrt(quad);
        } else if (c < 0.) {
          return -k + std::sqrt(quad);
        } else {
          const double d = -k - std::sqrt(quad);
          if (d < 0.)
            return INF;
          else
            return d;
        }
      }
    }
  }
};

//==========================================================================
// Cells (src/cell.cpp)
struct Problem;
struct Cell {
  uint32_t id;
  int fill_type;
  int fill;
  std::vector<int32_t> rpn;
  bool simple = true;
  bool vacuum_or_reflective = false;
};

struct Pad {  // include/geometry/geo_lily_pad.hpp
  enum Type { Universe = 0, Lattice = 1, CellT = 2 };
  int type;
  int idx;  // universe index (Universe/Lattice) or cell index (Cell)
  V3 r_local;
  int tile[3];
  bool in_outer;
};

struct BoundaryInfo {  // include/geometry/boundary.hpp
  double distance;
  int surface_index;
  int btype;
  int32_t token;
};

struct Universe {
  uint32_t id;
  int type;
  std::vector<int> cells;
  bool has_bc = false;
  // rect lattice (src/rect_lattice.cpp:33-52)
  uint32_t Nx = 0, Ny = 0, Nz = 0;
  double Px = 0, Py = 0, Pz = 0, Px_inv = 0, Py_inv = 0, Pz_inv = 0;
  double Xl = 0, Yl = 0, Zl = 0;
  std::vector<int32_t> elems;
  int outer = -1;
  // hex lattice (src/hex_lattice.cpp:30-56; include/geometry/hex_lattice.hpp:
  // 58-66): pitch in Px, axial pitch in Pz, origin X_o/Y_o/Z_o in Xl/Yl/Zl,
  // tiles in a width x width x Nz array of axial (q, r) shifted by mid
  uint32_t rings = 0, width = 0, mid = 0;
  int top = ABX_HEX_POINTY;
  double c6 = 0, s6 = 0, c3 = 0, s3 = 0;
};

//==========================================================================
// Multigroup nuclide (src/mg_nuclide.cpp) + angle tables
struct AngleDist {  // include/materials/mg_angle_distribution.hpp
  std::vector<double> mu{-1., 1.}, pdf{0.5, 0.5}, cdf{0., 1.};
  double sample_mu(Pcg32& g) const {
    const double xi = RNG::rand(g);
    auto it = std::lower_bound(cdf.begin(), cdf.end(), xi);
    std::size_t l = (std::size_t)std::distance(cdf.begin(), it);
    if (xi == *it) return mu[l];
    l--;
    if (pdf[l] == pdf[l + 1]) return mu[l] + ((xi - cdf[l]) / pdf[l]);
    double m = (pdf[l + 1] - pdf[l]) / (mu[l + 1] - mu[l]);
    return mu[l] +
           (1. / m) * (std::sqrt(pdf[l] * pdf[l] + 2. * m * (xi - cdf[l])) -
                       pdf[l]);
  }
};

struct Material {
  uint32_t id;
  int G;
  std::vector<double> Et, Ea, Ef, nu_p, nu_d, Es;
  std::vector<std::vector<double>> chi, Ps;
  std::vector<std::vector<AngleDist>> angle;
  std::vector<double> P_delayed_group, delayed_constants;
  bool fissile = false;
  bool has_nu_d = false;
};

struct MicroXS {  // include/materials/nuclide.hpp:40-53
  double total, fission, absorption, elastic, nu_total, nu_delayed;
  int energy_index;
};

struct MeshTally {
  int estimator, quantity;
  int Nx, Ny, Nz;
  V3 r_low, r_hi;
  double dx, dy, dz, dx_inv, dy_inv, dz_inv;
  double net_weight = 1.;
  std::vector<double> ebounds;
  std::vector<double> gen, avg, var;
  int g = 0;
};

struct ParticleState {
  V3 r, u;
  double E, w, w2;
};

struct Particle {  // include/simulation/particle.hpp:68-242
  ParticleState s;
  Pcg32 rng;
  uint64_t history_id = 0, family_id = 0, secondary_id = 0,
           daughter_counter = 0;
  V3 prev_r{0, 0, 0}, prev_u{1, 0, 0};
  double prev_E = 0, Esmp = 0;
  bool alive = true, reflected = false, prev_virtual = false;
  V3 r_birth;
  std::vector<ParticleState> secondaries;
  std::vector<abx_banked> fission;
  std::vector<abx_parent_info> fission_parent;  // beside `fission`

  void move(double d) {
    if (!reflected) {
      prev_r = s.r;
      s.r = add(s.r, scale(s.u, d));
    } else {
      s.r = add(s.r, scale(s.u, d));
      reflected = false;
    }
  }
  void set_position(V3 r) {
    prev_r = s.r;
    s.r = r;
  }
  void set_direction(V3 u) {
    prev_u = s.u;
    s.u = u;
  }
  void set_energy(double E) {
    prev_E = s.E;
    s.E = E;
  }
  void make_secondary(V3 u, double E, double w, double w2) {
    secondaries.push_back({s.r, u, E, w, w2});
  }
  void resurect() {
    if (!alive && !secondaries.empty()) {
      alive = true;
      s = secondaries.back();
      secondaries.pop_back();
      secondary_id++;
    }
  }
  void split(int n_new) {
    if (n_new > 1) {
      s.w = s.w / (double)n_new;
      s.w2 = s.w2 / (double)n_new;
      for (int np = 0; np < n_new - 1; np++) make_secondary(s.u, s.E, s.w, s.w2);
    }
  }
};

struct Scores {  // ThreadLocalScores (transporter.hpp:52-59)
  double k_col = 0, k_abs = 0, k_trk = 0, k_tot = 0, leak = 0, mig = 0;
};

struct Source {
  abx_source d;
};

//==========================================================================
struct Problem {
  std::vector<Surface> surfaces;
  std::vector<Cell> cells;
  std::vector<Universe> universes;
  int root = 0;
  std::vector<Material> materials;
  std::vector<Source> sources;
  std::vector<MeshTally> tallies;
  abx_settings st;
  std::vector<double> ebounds;  // settings::energy_bounds
  std::vector<double> ratio;    // sample_xs_ratio
  std::vector<double> majorant, sampling;
  double min_energy = 0., max_energy = 100000.;
  bool portable = true;

  // Tallies scalars (src/tallies.cpp)
  Scores gen_scores;
  double kcol_prev = 1.0;
  int gen = 0;
  double k_col_avg = 0, k_col_var = 0, k_abs_avg = 0, k_abs_var = 0, k_trk_avg = 0,
         k_trk_var = 0, k_tot_avg = 0, k_tot_var = 0, leak_avg = 0, leak_var = 0, mig_avg = 0,
         mig_var = 0;
  std::vector<abx_particle> last_bank;
  // Tallies::{k_col,k_abs,k_trk,k_tot,leak,mig}_vec: every generation's values
  std::vector<double> gen_vals[6];
  abx_diag diag_total{};
  // ApproximateMeshCancelator (src/approximate_mesh_cancelator.cpp:32-96)
  bool cancel = false;
  V3 c_low{0, 0, 0}, c_hi{0, 0, 0};
  uint32_t c_shape[4] = {1, 1, 1, 1};
  double c_dx = 0, c_dy = 0, c_dz = 0;
  std::vector<double> c_edges;
  // ExactMGCancelator / BasicExactMGCancelator, and settings::rng
  int exact = 0;  // abx_exact_kind, 0 = none
  abx_exact_cancel ex{};
  std::vector<std::vector<int>> ex_groups;
  V3 ex_low{0, 0, 0}, ex_hi{0, 0, 0};
  double ex_pitch[3] = {0, 0, 0};
  uint32_t ex_shape[4] = {1, 1, 1, 1};
  Pcg32 grng;
  std::vector<abx_parent_info> last_parents;  // beside the last transport's bank
  int universe_by_id(uint32_t id) const {
    for (size_t i = 0; i < universes.size(); i++)
      if (universes[i].id == id) return (int)i;
    return 0;  // std::map::operator[] default-inserts index 0
  }
};

//------------------------------------------------------------------------
// Cell::is_inside_simple / is_inside_complex (src/cell.cpp:144-193)
static bool cell_is_inside(const Problem& P, const Cell& c, const V3& r,
                           const V3& u, int32_t on_surf) {
  if (c.simple) {
    for (int32_t token : c.rpn) {
      if (token == on_surf) {
      } else if (-token == on_surf)
        return false;
      else {
        int sign = P.surfaces[(size_t)(std::abs(token) - 1)].sign(r, u);
        if ((sign > 0 && token < 0) || (sign < 0 && token > 0)) return false;
      }
    }
    return true;
  }
  std::vector<bool> stck(c.rpn.size());
  int i = -1;
  for (int32_t token : c.rpn) {
    if (token == ABX_OP_UNIN) {
      stck[(size_t)(i - 1)] = stck[(size_t)(i - 1)] || stck[(size_t)i];
      i--;
    } else if (token == ABX_OP_INTR) {
      stck[(size_t)(i - 1)] = stck[(size_t)(i - 1)] && stck[(size_t)i];
      i--;
    } else if (token == ABX_OP_COMP) {
      stck[(size_t)i] = !stck[(size_t)i];
    } else {
      i++;
      if (token == on_surf)
        stck[(size_t)i] = true;
      else if (-token == on_surf)
        stck[(size_t)i] = false;
      else {
        int sign = P.surfaces[(size_t)(std::abs(token) - 1)].sign(r, u);
        stck[(size_t)i] = (sign > 0 && token > 0) || (sign < 0 && token < 0);
      }
    }
  }
  if (i == 0) return stck[0];
  return true;
}

// Cell::distance_to_boundary (src/cell.cpp:79-106)
static std::pair<double, int32_t> cell_distance(const Problem& P,
                                                const Cell& c, const V3& r,
                                                const V3& u, int32_t on_surf,
                                                bool bc_only) {
  if (bc_only && !c.vacuum_or_reflective) return {INF, 0};
  double min_dist = INF;
  int32_t i_surf = 0;
  for (int32_t token : c.rpn) {
    if (token >= ABX_OP_UNIN) continue;
    bool coincident = std::abs(token) == std::abs(on_surf);
    const Surface& s = P.surfaces[(size_t)(std::abs(token) - 1)];
    if (bc_only && s.bc == ABX_NORMAL) continue;
    double d = s.distance(r, u, coincident);
    if (d < min_dist) {
      if (std::abs(d - min_dist) / min_dist >= 1e-14) {
        min_dist = d;
        i_surf = -token;
      }
    }
  }
  return {min_dist, i_surf};
}

// RectLattice::get_tile (src/rect_lattice.cpp:209-237)
static void lattice_tile_center(const Universe& L, int nx, int ny, int nz,
                                V3& c) {
  c.x = ((double)nx + 0.5) * L.Px + L.Xl;
  c.y = ((double)ny + 0.5) * L.Py + L.Yl;
  c.z = ((double)nz + 0.5) * L.Pz + L.Zl;
}

static void lattice_get_tile(const Universe& L, const V3& r, const V3& u,
                             int t[3]) {
  int nx = (int)std::floor((r.x - L.Xl) * L.Px_inv);
  int ny = (int)std::floor((r.y - L.Yl) * L.Py_inv);
  int nz = (int)std::floor((r.z - L.Zl) * L.Pz_inv);
  V3 c;
  lattice_tile_center(L, nx, ny, nz, c);
  double xl = c.x - L.Px * 0.5;
  if (std::abs(xl - r.x) < SURFACE_COINCIDENT && u.x < 0.) nx--;
  double xh = c.x + L.Px * 0.5;
  if (std::abs(xh - r.x) < SURFACE_COINCIDENT && u.x >= 0.) nx++;
  double yl = c.y - L.Py * 0.5;
  if (std::abs(yl - r.y) < SURFACE_COINCIDENT && u.y < 0.) ny--;
  double yh = c.y + L.Py * 0.5;
  if (std::abs(yh - r.y) < SURFACE_COINCIDENT && u.y >= 0.) ny++;
  double zl = c.z - L.Pz * 0.5;
  if (std::abs(zl - r.z) < SURFACE_COINCIDENT && u.z < 0.) nz--;
  double zh = c.z + L.Pz * 0.5;
  if (std::abs(zh - r.z) < SURFACE_COINCIDENT && u.z >= 0.) nz++;
  t[0] = nx;
  t[1] = ny;
  t[2] = nz;
}

// RectLattice::distance_to_tile_boundary (src/rect_lattice.cpp:239-282)
static double lattice_tile_distance(const Universe& L, const V3& r_local,
                                    const V3& u, const int tile[3]) {
  V3 c;
  lattice_tile_center(L, tile[0], tile[1], tile[2], c);
  V3 rt = sub(r_local, c);
  double dist = INF;
  const double diff_xl = -L.Px * 0.5 - rt.x;
  const double diff_xh = L.Px * 0.5 - rt.x;
  const double diff_yl = -L.Py * 0.5 - rt.y;
  const double diff_yh = L.Py * 0.5 - rt.y;
  const double diff_zl = -L.Pz * 0.5 - rt.z;
  const double diff_zh = L.Pz * 0.5 - rt.z;
  const double ux_inv = 1. / u.x;
  const double uy_inv = 1. / u.y;
  const double uz_inv = 1. / u.z;
  const double d_xl = diff_xl * ux_inv;
  const double d_xh = diff_xh * ux_inv;
  const double d_yl = diff_yl * uy_inv;
  const double d_yh = diff_yh * uy_inv;
  const double d_zl = diff_zl * uz_inv;
  const double d_zh = diff_zh * uz_inv;
  const double tol = 100 * SURFACE_COINCIDENT;
  if (d_xl > 0. && d_xl < dist && std::abs(diff_xl) > tol) dist = d_xl;
  if (d_xh > 0. && d_xh < dist && std::abs(diff_xh) > tol) dist = d_xh;
  if (d_yl > 0. && d_yl < dist && std::abs(diff_yl) > tol) dist = d_yl;
  if (d_yh > 0. && d_yh < dist && std::abs(diff_yh) > tol) dist = d_yh;
  if (d_zl > 0. && d_zl < dist && std::abs(diff_zl) > tol) dist = d_zl;
  if (d_zh > 0. && d_zh < dist && std::abs(diff_zh) > tol) dist = d_zh;
  return dist;
}

static bool tile_in_range(const Universe& L, const int t[3]) {
  return !(t[0] < 0 || t[0] >= (int)L.Nx || t[1] < 0 || t[1] >= (int)L.Ny ||
           t[2] < 0 || t[2] >= (int)L.Nz);
}

static size_t lattice_linear_index(const Universe& L, int nx, int ny, int nz) {
  return (size_t)((uint32_t)nz * (L.Nx * L.Ny) + (uint32_t)nx * L.Ny +
                  (uint32_t)ny);
}

//--------------------------------------------------------------------------
// HexLattice (src/hex_lattice.cpp).  The arithmetic keeps the reference's
// expressions term by term (its 0.* and 1.* factors included), so that the
// tiles and distances carry the same bits.

// HexLattice::get_nearest_hex (src/hex_lattice.cpp:260-301): axial (q, r)
// of the point in the lattice's centred frame, by cube rounding.
static void hex_nearest(const Universe& L, const V3& p, int32_t qr[2]) {
  const double pt = L.Px;
  double q, r, det;
  if (L.top == ABX_HEX_FLAT) {
    det = -pt * pt * L.s3;
    q = (pt / det) * (0. * p.x - 1. * p.y);
    r = (pt / det) * (-L.s3 * p.x + L.c3 * p.y);
  } else {
    det = -pt * pt * L.c6;
    q = (pt / det) * (L.s6 * p.x - L.c6 * p.y);
    r = (pt / det) * (-1. * p.x - 0. * p.y);
  }
  const double cx = q, cz = r, cy = -cx - cz;
  double rx = std::round(cx), ry = std::round(cy), rz = std::round(cz);
  const double dx = std::abs(rx - cx), dy = std::abs(ry - cy), dz = std::abs(rz - cz);
  if (dx > dy && dx > dz)
    rx = -ry - rz;
  else if (dy > dx && dy > dz)
    ry = -rx - rz;
  else
    rz = -rx - ry;
  qr[0] = (int32_t)rx;
  qr[1] = (int32_t)rz;
}

// HexLattice::get_tile (src/hex_lattice.cpp:303-308): the direction plays
// no part, and the axial layer is counted from Z_o - Nz pz / 2 in the frame
// the caller passes.
static void hex_get_tile(const Universe& L, const V3& p, int t[3]) {
  int32_t qr[2];
  hex_nearest(L, p, qr);
  const double z_low = L.Zl - 0.5 * (double)L.Nz * L.Pz;
  t[0] = qr[0];
  t[1] = qr[1];
  t[2] = (int32_t)std::floor((p.z - z_low) / L.Pz);
}

// HexLattice::tile_center / get_hex_center (src/hex_lattice.cpp:310-331)
static V3 hex_tile_center(const Universe& L, int q, int r, int nz) {
  const double z_low = L.Zl - 0.5 * (double)L.Nz * L.Pz;
  const double dq = (double)q, dr = (double)r;
  V3 c;
  if (L.top == ABX_HEX_FLAT) {
    c.x = L.Px * (L.c3 * dq + 1. * dr);
    c.y = L.Px * (L.s3 * dq + 0. * dr);
  } else {
    c.x = L.Px * (0. * dq + L.c6 * dr);
    c.y = L.Px * (1. * dq + L.s6 * dr);
  }
  c.z = ((double)nz + 0.5) * L.Pz + z_low;
  return c;
}

// HexLattice::get_ring (src/hex_lattice.cpp:333-345)
static uint32_t hex_ring(int32_t q, int32_t r) {
  const int32_t y = -q - r;
  return std::max(std::max((uint32_t)std::abs(q), (uint32_t)std::abs(y)),
                  (uint32_t)std::abs(r));
}

// HexLattice::distance_to_line (src/hex_lattice.cpp:435-452)
static double hex_line(const V3& r, const V3& u, double x1, double y1, double x2,
                       double y2) {
  const double A = y2 - y1;
  const double B = x1 - x2;
  const double D = (x2 - x1) * y1 - (y2 - y1) * x1;
  const double num = D - A * r.x - B * r.y;
  const double denom = A * u.x + B * u.y;
  const double d = num / denom;
  if (d < 0.) return HUGE_VAL;
  // within the coincidence tolerance of the line (|num| / |(A, B)|)
  if (std::abs(num) <= 100 * SURFACE_COINCIDENT * std::sqrt(A * A + B * B)) return HUGE_VAL;
  return d;
}

// HexLattice::distance_to_tile_boundary and its pointy / flat parts
// (src/hex_lattice.cpp:347-433): six edge lines and two axial planes in the
// frame of the tile centre, taken from r_local as given.  The six lines are
// the reference's, vertex values included; they are not the edges of the
// hexagon get_nearest_hex assigns (a particle can stop short of, or run
// past, its tile's true edge) and are kept as they are for equal results.
//
// One deviation.  The reference takes any positive distance, and its tiles
// do not depend on the direction.  None of its six lines is an edge of the
// tile, so crossing one keeps the tile: a particle stopped on a line finds
// it again ~1e-16 cm ahead (half the time), too close to move it, and
// SurfaceTracker::transport (src/surface_tracker.cpp:72-150) crosses it for
// ever (the first histories of a surface-tracked hex lattice already do).
// Lines and layer planes within 100 * SURFACE_COINCIDENT (1e-10 cm) of the
// point are skipped here, the tolerance RectLattice applies to its own
// planes (src/rect_lattice.cpp:268-281).  A history the reference completes
// differs only if a flight starts within 1e-10 cm of a line or plane
// without stopping on it.
static double hex_tile_distance(const Universe& L, const V3& r_local, const V3& u,
                                const int tile[3]) {
  const V3 rt = sub(r_local, hex_tile_center(L, tile[0], tile[1], tile[2]));
  const double pt = L.Px;
  double e[6];
  if (L.top == ABX_HEX_POINTY) {
    const double x1 = 0., y1 = pt / (2. * L.c6), x2 = pt / 2., y2 = pt / (2. * L.s6);
    e[0] = hex_line(rt, u, x1, y1, x2, y1);
    e[1] = hex_line(rt, u, x2, y2, x2, -y2);
    e[2] = hex_line(rt, u, x2, -y2, x1, -y1);
    e[3] = hex_line(rt, u, x1, -y1, -x2, -y2);
    e[4] = hex_line(rt, u, -x2, -y2, -x2, y2);
    e[5] = hex_line(rt, u, -x2, y2, x1, y1);
  } else {
    const double x1 = pt / (2. * L.s6), y1 = pt / 2., x2 = pt / (2. * L.c6), y2 = 0.;
    e[0] = hex_line(rt, u, x1, y1, x2, y1);
    e[1] = hex_line(rt, u, x2, y2, x1, -y1);
    e[2] = hex_line(rt, u, x1, -y1, -x1, -y1);
    e[3] = hex_line(rt, u, -x1, -y1, -x2, -y2);
    e[4] = hex_line(rt, u, -x2, -y2, -x1, y1);
    e[5] = hex_line(rt, u, -x1, y1, x1, y1);
  }
  const double dzl = (-L.Pz * 0.5 - rt.z) / u.z;
  const double dzu = (L.Pz * 0.5 - rt.z) / u.z;
  const double tol = 100 * SURFACE_COINCIDENT;
  double d = INF;
  for (int k = 0; k < 6; k++)
    if (e[k] > 0. && e[k] < d) d = e[k];
  if (dzl > 0. && dzl < d && std::abs(-L.Pz * 0.5 - rt.z) > tol) d = dzl;
  if (dzu > 0. && dzu < d && std::abs(L.Pz * 0.5 - rt.z) > tol) d = dzu;
  return d;
}

// The element of a hex tile, -1 outside the rings or the layers
// (src/hex_lattice.cpp:143-206, 245-258).
static int hex_element(const Universe& L, const int t[3]) {
  if (hex_ring(t[0], t[1]) >= L.rings) return -1;
  if (t[2] < 0 || t[2] >= (int32_t)L.Nz) return -1;
  const size_t li = (size_t)((uint32_t)t[2] * (L.width * L.width) +
                             ((uint32_t)t[1] + L.mid) * L.width + ((uint32_t)t[0] + L.mid));
  return L.elems[li];
}

// Lattice::get_tile / distance_to_tile_boundary for either lattice kind,
// from a lattice pad's r_local (the frame the lattice was entered in).
static void lattice_pad_tile(const Universe& L, const V3& r, const V3& u, int t[3]) {
  if (L.type == ABX_HEX_LATTICE)
    hex_get_tile(L, r, t);
  else
    lattice_get_tile(L, r, u, t);
}
static double lattice_pad_distance(const Universe& L, const V3& r, const V3& u,
                                   const int tile[3]) {
  return L.type == ABX_HEX_LATTICE ? hex_tile_distance(L, r, u, tile)
                                   : lattice_tile_distance(L, r, u, tile);
}

// Universe::get_cell with the lily-pad stack: CellUniverse
// (src/cell_universe.cpp:72-109) and RectLattice (src/rect_lattice.cpp:128-207).
// Returns the material cell index or -1.
static int universe_get_cell(const Problem& P, int ui, std::vector<Pad>& stack,
                             V3 r, V3 u, int32_t on_surf) {
  const Universe& U = P.universes[(size_t)ui];
  if (U.type == ABX_CELL_UNIVERSE) {
    stack.push_back({Pad::Universe, ui, r, {0, 0, 0}, false});
    for (int ci : U.cells) {
      const Cell& c = P.cells[(size_t)ci];
      if (cell_is_inside(P, c, r, u, on_surf)) {
        stack.push_back({Pad::CellT, ci, r, {0, 0, 0}, false});
        if (c.fill_type == ABX_FILL_MATERIAL) return ci;
        return universe_get_cell(P, c.fill, stack, r, u, on_surf);
      }
    }
    return -1;
  }
  if (U.type == ABX_HEX_LATTICE) {
    // HexLattice::get_cell (src/hex_lattice.cpp:143-217): the tile of the
    // origin-relative point; the pad keeps r, the element's frame is the
    // origin-relative point less the tile centre, and the outer universe
    // (if any) is entered at r.
    const V3 ro{r.x - U.Xl, r.y - U.Yl, r.z - U.Zl};
    int t[3];
    hex_get_tile(U, ro, t);
    const int e = hex_element(U, t);
    if (e >= 0) {
      const V3 c = hex_tile_center(U, t[0], t[1], t[2]);
      stack.push_back({Pad::Lattice, ui, r, {t[0], t[1], t[2]}, false});
      return universe_get_cell(P, e, stack, sub(ro, c), u, on_surf);
    }
    stack.push_back({Pad::Lattice, ui, r, {t[0], t[1], t[2]}, U.outer >= 0});
    if (U.outer >= 0) return universe_get_cell(P, U.outer, stack, r, u, on_surf);
    return -1;
  }
  int t[3];
  lattice_get_tile(U, r, u, t);
  if (!tile_in_range(U, t)) {
    if (U.outer >= 0) {
      stack.push_back({Pad::Lattice, ui, r, {t[0], t[1], t[2]}, true});
      return universe_get_cell(P, U.outer, stack, r, u, on_surf);
    }
    stack.push_back({Pad::Lattice, ui, r, {t[0], t[1], t[2]}, false});
    return -1;
  }
  size_t li = lattice_linear_index(U, t[0], t[1], t[2]);
  if (U.elems[li] >= 0) {
    V3 c;
    lattice_tile_center(U, t[0], t[1], t[2], c);
    V3 r_local = sub(r, c);
    stack.push_back({Pad::Lattice, ui, r, {t[0], t[1], t[2]}, false});
    return universe_get_cell(P, U.elems[li], stack, r_local, u, on_surf);
  }
  if (U.outer >= 0) {
    stack.push_back({Pad::Lattice, ui, r, {t[0], t[1], t[2]}, true});
    return universe_get_cell(P, U.outer, stack, r, u, on_surf);
  }
  stack.push_back({Pad::Lattice, ui, r, {t[0], t[1], t[2]}, false});
  return -1;
}

// Universe::get_boundary_condition: CellUniverse (src/cell_universe.cpp:111-150)
// and Lattice (src/lattice.cpp:77-91).
static BoundaryInfo universe_boundary_condition(const Problem& P, int ui,
                                                const V3& r, const V3& u,
                                                int32_t on_surf) {
  const Universe& U = P.universes[(size_t)ui];
  if (U.type != ABX_CELL_UNIVERSE) {
    if (U.has_bc) return universe_boundary_condition(P, U.outer, r, u, on_surf);
    return {INF, -1, ABX_VACUUM, 0};
  }
  double dist = INF;
  int btype = ABX_VACUUM;
  int surface_index = -1;
  int32_t token = 0;
  if (U.has_bc) {
    for (int ci : U.cells) {
      const Cell& c = P.cells[(size_t)ci];
      if (!c.vacuum_or_reflective) continue;
      auto dt = cell_distance(P, c, r, u, on_surf, true);
      if (dt.first < dist && std::abs(dt.first - dist) > BOUNDRY_TOL) {
        double tmp_dist = dt.first;
        int32_t tmp_token = std::abs(dt.second);
        if (tmp_token) {
          token = tmp_token;
          dist = tmp_dist;
          surface_index = token - 1;
        } else {
          continue;
        }
        btype = P.surfaces[(size_t)surface_index].bc;
        if (P.surfaces[(size_t)surface_index].sign(r, u) < 0) token *= -1;
      }
    }
  }
  return {dist, surface_index, btype, token};
}

//==========================================================================
// Tracker (include/simulation/tracker.hpp:41-372)
struct Tracker {
  const Problem* P;
  V3 r_, u_;
  std::vector<Pad> tree;
  int cell = -1;
  int32_t token = 0;

  Tracker(const Problem* p, V3 r, V3 u, int32_t tok = 0)
      : P(p), r_(r), u_(u), token(tok) {
    tree.reserve(10);
    cell = universe_get_cell(*P, P->root, tree, r_, u_, token);
  }
  bool is_lost() const { return cell < 0; }
  int material() const {
    return cell < 0 ? -1 : P->cells[(size_t)cell].fill;
  }
  void set_r(V3 r) {
    r_ = r;
    token = 0;
  }
  void set_u(V3 u) { u_ = u; }
  void restart_get_current() {
    tree.clear();
    cell = universe_get_cell(*P, P->root, tree, r_, u_, token);
  }
  void move(double d) {
    r_ = add(r_, scale(u_, d));
    for (auto& leaf : tree) leaf.r_local = add(leaf.r_local, scale(u_, d));
    token = 0;
  }
  bool check_tree() const {
    return r_.x == tree.front().r_local.x && r_.y == tree.front().r_local.y &&
           r_.z == tree.front().r_local.z;
  }
  // Tracker::get_current (tracker.hpp:235-306)
  void get_current() {
    size_t first_bad = tree.size();
    if (!check_tree()) {
      restart_get_current();
      if (!check_tree()) fatal(ABX_E_LOST, "BAD POSITIONS");
    }
    for (size_t i = 0; i < tree.size(); i++) {
      const Pad& pad = tree[i];
      if (pad.type == Pad::CellT) {
        if (!cell_is_inside(*P, P->cells[(size_t)pad.idx], pad.r_local, u_,
                            token)) {
          first_bad = i;
          break;
        }
      } else if (pad.type == Pad::Lattice) {
        int t[3];
        lattice_pad_tile(P->universes[(size_t)pad.idx], pad.r_local, u_, t);
        if (pad.tile[0] != t[0] || pad.tile[1] != t[1] || pad.tile[2] != t[2]) {
          first_bad = i;
          break;
        }
      }
    }
    if (first_bad != tree.size()) {
      if (first_bad == 0) return restart_get_current();
      // The reference keeps an end() iterator of the tree from before a
      // restart; when the restarted tree is shallower it resizes into
      // default-constructed pads.  That only happens after losing a particle
      // deeper than where it is found again (a geometry hole); report it.
      if (first_bad > tree.size())
        fatal(ABX_E_LOST, "relocation from a deeper lost tree");
      tree.resize(first_bad);
      // universe_id_to_indx[tree.back().id]: a Cell pad's id is looked up
      // among universe ids (tracker.hpp:285-286).
      const Pad& back = tree.back();
      int ui = back.type == Pad::CellT
                   ? P->universe_by_id(P->cells[(size_t)back.idx].id)
                   : back.idx;
      V3 r_local = back.r_local;
      tree.pop_back();
      cell = universe_get_cell(*P, ui, tree, r_local, u_, token);
      if (cell < 0) restart_get_current();
      if (cell >= 0 && P->cells[(size_t)cell].fill_type != ABX_FILL_MATERIAL)
        fatal(ABX_E_LOST, "Did not find a cell with a material.");
    }
  }
  // Tracker::get_boundary_condition (tracker.hpp:94-161)
  BoundaryInfo get_boundary_condition() const {
    if (is_lost())
      return universe_boundary_condition(*P, P->root, r_, u_, token);
    double dist = INF;
    int btype = ABX_VACUUM;
    int surface_index = -1;
    int32_t tok = 0;
    for (const Pad& pad : tree) {
      if (pad.type == Pad::CellT) {
        const Cell& c = P->cells[(size_t)pad.idx];
        if (!c.vacuum_or_reflective) continue;
        auto dt = cell_distance(*P, c, pad.r_local, u_, token, true);
        if (dt.first < dist && std::abs(dt.first - dist) > BOUNDRY_TOL) {
          double tmp_dist = dt.first;
          int32_t tmp_token = std::abs(dt.second);
          if (tmp_token) {
            dist = tmp_dist;
            tok = tmp_token;
            surface_index = tok - 1;
          } else {
            continue;
          }
          btype = P->surfaces[(size_t)surface_index].bc;
          if (P->surfaces[(size_t)surface_index].sign(pad.r_local, u_) < 0)
            tok *= -1;
        }
      } else {
        const Universe& U = P->universes[(size_t)pad.idx];
        if (U.has_bc) {
          BoundaryInfo ub =
              universe_boundary_condition(*P, pad.idx, pad.r_local, u_, token);
          if (ub.distance < dist && std::abs(ub.distance - dist) > BOUNDRY_TOL) {
            dist = ub.distance;
            tok = ub.token;
            surface_index = ub.surface_index;
            btype = ub.btype;
          }
        }
      }
    }
    return {dist, surface_index, btype, tok};
  }
  // Tracker::get_nearest_boundary (tracker.hpp:163-225)
  BoundaryInfo get_nearest_boundary() const {
    if (is_lost()) fatal(ABX_E_LOST, "nearest boundary of a lost particle");
    BoundaryInfo b = get_boundary_condition();
    double dist = b.distance;
    int btype = b.btype;
    int surface_index = b.surface_index;
    int32_t tok = b.token;
    for (const Pad& pad : tree) {
      if (pad.type == Pad::Lattice) {
        const Universe& L = P->universes[(size_t)pad.idx];
        double d = lattice_pad_distance(L, pad.r_local, u_, pad.tile);
        if (d < dist && std::abs(d - dist) > BOUNDRY_TOL) {
          dist = d;
          btype = ABX_NORMAL;
          surface_index = -1;
          tok = 0;
        }
      } else if (pad.type == Pad::CellT) {
        auto dt = cell_distance(*P, P->cells[(size_t)pad.idx], pad.r_local, u_,
                                token, false);
        if (dt.first < dist && std::abs(dt.first - dist) > BOUNDRY_TOL) {
          dist = dt.first;
          tok = std::abs(dt.second);
          surface_index = tok ? tok - 1 : -1;
          if (surface_index >= 0)
            btype = P->surfaces[(size_t)surface_index].bc;
          else
            btype = ABX_NORMAL;
          if (surface_index >= 0 &&
              P->surfaces[(size_t)surface_index].sign(pad.r_local, u_) < 0)
            tok *= -1;
        }
      }
    }
    return {dist, surface_index, btype, tok};
  }
  void cross_surface(const BoundaryInfo& b) {
    move(b.distance);
    token = -b.token;
  }
  // Tracker::do_reflection (tracker.hpp:314-359)
  void do_reflection(Particle& p, const BoundaryInfo& b) {
    if (b.surface_index < 0) fatal(ABX_E_LOST, "Bad surface index in reflection");
    const Surface& s = P->surfaces[(size_t)b.surface_index];
    V3 r_pre_refs = p.s.r;
    if (p.reflected) r_pre_refs = p.prev_r;
    V3 u = p.s.u;
    V3 r_on_surf = add(p.s.r, scale(u, b.distance));
    V3 n = s.norm(r_on_surf);
    V3 nd = sub(u, scale(n, 2. * dot(u, n)));
    V3 u_new = direction(nd.x, nd.y, nd.z);
    double d = b.distance + norm(sub(p.s.r, r_pre_refs));
    V3 r_prev = sub(r_on_surf, scale(u_new, d));
    p.set_position(r_on_surf);
    p.prev_r = r_prev;
    p.set_direction(u_new);
    p.reflected = true;
    set_r(p.s.r);
    set_u(p.s.u);
    restart_get_current();
  }
};

//==========================================================================
// Material helpers (include/materials/material_helper.hpp, MG nuclide xs
// getters src/mg_nuclide.cpp:348-411)
static int energy_grid_index(const Problem& P, double E) {
  size_t i = 0;
  for (i = 0; i < P.ebounds.size() - 1; i++)
    if (P.ebounds[i] <= E && E < P.ebounds[i + 1]) break;
  return (int)i;
}

static MicroXS micro_xs(const Problem& P, const Material& m, double E) {
  MicroXS x;
  x.energy_index = energy_grid_index(P, E);
  if (x.energy_index >= m.G) fatal(ABX_E_INVALID, "energy outside group bounds");
  const size_t i = (size_t)x.energy_index;
  x.total = m.Et[i];
  x.fission = m.Ef[i];
  x.absorption = x.fission + (m.Ea[i] - m.Ef[i]);
  x.elastic = m.Es[i];
  x.nu_total = m.nu_p[i];
  if (m.has_nu_d) x.nu_total += m.nu_d[i];
  x.nu_delayed = m.has_nu_d ? m.nu_d[i] : 0.;
  return x;
}

// macroscopic sums over the single component at 1 atom/b-cm
static double mat_Et(const Problem& P, int mi, double E) {
  return 0. + 1.0 * micro_xs(P, P.materials[(size_t)mi], E).total;
}
static double mat_Ea(const Problem& P, int mi, double E) {
  return 0. + 1.0 * micro_xs(P, P.materials[(size_t)mi], E).absorption;
}
static double mat_Ef(const Problem& P, int mi, double E) {
  return 0. + 1.0 * micro_xs(P, P.materials[(size_t)mi], E).fission;
}
static double mat_vEf(const Problem& P, int mi, double E) {
  MicroXS x = micro_xs(P, P.materials[(size_t)mi], E);
  return 0. + 1.0 * x.nu_total * x.fission;
}
static double mat_Eel(const Problem& P, int mi, double E) {
  return 0. + 1.0 * micro_xs(P, P.materials[(size_t)mi], E).elastic;
}

//==========================================================================
// Tallies (src/collision_mesh_tally.cpp:33-110, src/track_length_mesh_tally.cpp)
static double quantity_factor(const Problem& P, const MeshTally& t,
                              const Particle& p, int mi) {
  switch (t.quantity) {
    case ABX_Q_ELASTIC:
      return p.s.w * mat_Eel(P, mi, p.s.E);
    case ABX_Q_ABSORPTION:
      return p.s.w * mat_Ea(P, mi, p.s.E);
    case ABX_Q_FISSION:
      return p.s.w * mat_Ef(P, mi, p.s.E);
    case ABX_Q_TOTAL:
      return p.s.w * mat_Et(P, mi, p.s.E);
    case ABX_Q_IMG_FLUX:
      return p.s.w2;
    default:
      return p.s.w;
  }
}

static int tally_ebin(const MeshTally& t, double E) {
  for (size_t e = 0; e < t.ebounds.size() - 1; e++)
    if (t.ebounds[e] <= E && E <= t.ebounds[e + 1]) return (int)e;
  return -1;
}

static inline void atomic_add(double& a, double v) {
#pragma omp atomic
  a += v;
}

static uint64_t score_collision(Problem& P, MeshTally& t, const Particle& p,
                                int mi) {
  double Et = mat_Et(P, mi, p.s.E);
  double scr = 1. / (Et * t.net_weight);
  int i = (int)std::floor((p.s.r.x - t.r_low.x) * t.dx_inv);
  int j = (int)std::floor((p.s.r.y - t.r_low.y) * t.dy_inv);
  int k = (int)std::floor((p.s.r.z - t.r_low.z) * t.dz_inv);
  int l = tally_ebin(t, p.s.E);
  if (l == -1) return 0;
  if (i >= 0 && i < t.Nx && j >= 0 && j < t.Ny && k >= 0 && k < t.Nz) {
    if (t.quantity == ABX_Q_TOTAL)
      scr *= p.s.w * Et;
    else
      scr *= quantity_factor(P, t, p, mi);
    size_t idx = (((size_t)l * t.Nx + (size_t)i) * t.Ny + (size_t)j) * t.Nz +
                 (size_t)k;
    atomic_add(t.gen[idx], scr);
    return 1;
  }
  return 0;
}

// TrackLengthMeshTally helpers (src/track_length_mesh_tally.cpp:182-393)
static bool tle_find_entry_point(const MeshTally& t, V3& r, const V3& u,
                                 double& d_flight) {
  const double ux_inv = 1. / u.x;
  const double uy_inv = 1. / u.y;
  const double uz_inv = 1. / u.z;
  double d_min = (t.r_low.x - r.x) * ux_inv;
  double d_max = (t.r_hi.x - r.x) * ux_inv;
  if (d_min > d_max) std::swap(d_min, d_max);
  double d_y_min = (t.r_low.y - r.y) * uy_inv;
  double d_y_max = (t.r_hi.y - r.y) * uy_inv;
  if (d_y_min > d_y_max) std::swap(d_y_min, d_y_max);
  if ((d_min > d_y_max) || (d_y_min > d_max)) return false;
  if (d_y_min > d_min) d_min = d_y_min;
  if (d_y_max < d_max) d_max = d_y_max;
  double d_z_min = (t.r_low.z - r.z) * uz_inv;
  double d_z_max = (t.r_hi.z - r.z) * uz_inv;
  if (d_z_min > d_z_max) std::swap(d_z_min, d_z_max);
  if ((d_min > d_z_max) || (d_z_min > d_max)) return false;
  if (d_z_min > d_min) d_min = d_z_min;
  if (d_z_max < d_max) d_max = d_z_max;
  if (d_max < d_min) std::swap(d_max, d_min);
  if ((d_max < 0.) && (d_min < 0.)) return false;
  if (d_min < 0.) return false;
  r = add(r, scale(u, d_min));
  d_flight -= d_min;
  return true;
}

static void tle_init_indices(const MeshTally& t, const V3& r, const V3& u,
                             int& i, int& j, int& k, int on[3]) {
  i = (int)std::floor((r.x - t.r_low.x) * t.dx_inv);
  j = (int)std::floor((r.y - t.r_low.y) * t.dy_inv);
  k = (int)std::floor((r.z - t.r_low.z) * t.dz_inv);
  on[0] = on[1] = on[2] = 0;
  double xc = t.r_low.x + i * t.dx + 0.5 * t.dx;
  double yc = t.r_low.y + j * t.dy + 0.5 * t.dy;
  double zc = t.r_low.z + k * t.dz + 0.5 * t.dz;
  double xl = xc - 0.5 * t.dx, xh = xc + 0.5 * t.dx;
  double yl = yc - 0.5 * t.dy, yh = yc + 0.5 * t.dy;
  double zl = zc - 0.5 * t.dz, zh = zc + 0.5 * t.dz;
  if (std::abs(xl - r.x) < SURFACE_COINCIDENT) {
    if (u.x < 0.) {
      i--;
      on[0] = 1;
    } else
      on[0] = -1;
  } else if (std::abs(xh - r.x) < SURFACE_COINCIDENT) {
    if (u.x < 0.)
      on[0] = 1;
    else {
      i++;
      on[0] = -1;
    }
  }
  if (std::abs(yl - r.y) < SURFACE_COINCIDENT) {
    if (u.y < 0.) {
      j--;
      on[1] = 1;
    } else
      on[1] = -1;
  } else if (std::abs(yh - r.y) < SURFACE_COINCIDENT) {
    if (u.y < 0.)
      on[1] = 1;
    else {
      j++;
      on[1] = -1;
    }
  }
  if (std::abs(zl - r.z) < SURFACE_COINCIDENT) {
    if (u.z < 0.) {
      k--;
      on[2] = 1;
    } else
      on[2] = -1;
  } else if (std::abs(zh - r.z) < SURFACE_COINCIDENT) {
    if (u.z < 0.)
      on[2] = 1;
    else {
      k++;
      on[2] = -1;
    }
  }
}

static std::pair<double, int> tle_next(const MeshTally& t, const V3& r,
                                       const V3& u, int i, int j, int k,
                                       const int on[3]) {
  double xc = t.r_low.x + i * t.dx + 0.5 * t.dx;
  double yc = t.r_low.y + j * t.dy + 0.5 * t.dy;
  double zc = t.r_low.z + k * t.dz + 0.5 * t.dz;
  V3 rt{r.x - xc, r.y - yc, r.z - zc};
  double dist = INF;
  int key = 0;
  const double diff_xl = -t.dx * 0.5 - rt.x;
  const double diff_xh = t.dx * 0.5 - rt.x;
  const double diff_yl = -t.dy * 0.5 - rt.y;
  const double diff_yh = t.dy * 0.5 - rt.y;
  const double diff_zl = -t.dz * 0.5 - rt.z;
  const double diff_zh = t.dz * 0.5 - rt.z;
  const double ux_inv = 1. / u.x;
  const double uy_inv = 1. / u.y;
  const double uz_inv = 1. / u.z;
  const double d_xl = diff_xl * ux_inv;
  const double d_xh = diff_xh * ux_inv;
  const double d_yl = diff_yl * uy_inv;
  const double d_yh = diff_yh * uy_inv;
  const double d_zl = diff_zl * uz_inv;
  const double d_zh = diff_zh * uz_inv;
  if (d_xl > 0. && d_xl < dist && on[0] != -1) {
    dist = d_xl;
    key = -1;
  }
  if (d_xh > 0. && d_xh < dist && on[0] != 1) {
    dist = d_xh;
    key = 1;
  }
  if (d_yl > 0. && d_yl < dist && on[1] != -1) {
    dist = d_yl;
    key = -2;
  }
  if (d_yh > 0. && d_yh < dist && on[1] != 1) {
    dist = d_yh;
    key = 2;
  }
  if (d_zl > 0. && d_zl < dist && on[2] != -1) {
    dist = d_zl;
    key = -3;
  }
  if (d_zh > 0. && d_zh < dist && on[2] != 1) {
    dist = d_zh;
    key = 3;
  }
  return {dist, key};
}

static uint64_t score_flight(Problem& P, MeshTally& t, const Particle& p,
                             double d, int mi) {
  V3 r = p.s.r;
  V3 u = p.s.u;
  int i = 0, j = 0, k = 0;
  int on[3] = {0, 0, 0};
  uint64_t n = 0;
  tle_init_indices(t, r, u, i, j, k, on);
  bool inside = i >= 0 && i < t.Nx && j >= 0 && j < t.Ny && k >= 0 && k < t.Nz;
  if (!inside) {
    if (!tle_find_entry_point(t, r, u, d)) return 0;
    tle_init_indices(t, r, u, i, j, k, on);
  }
  double base = 1. / t.net_weight;
  base *= quantity_factor(P, t, p, mi);
  int l = tally_ebin(t, p.s.E);
  if (l == -1) return 0;
  double remaining = d;
  while (remaining > 0.) {
    auto nt = tle_next(t, r, u, i, j, k, on);
    if (nt.first == INF) break;
    double d_tile = std::min(nt.first, remaining);
    if (i >= 0 && i < t.Nx && j >= 0 && j < t.Ny && k >= 0 && k < t.Nz) {
      size_t idx = (((size_t)l * t.Nx + (size_t)i) * t.Ny + (size_t)j) * t.Nz +
                   (size_t)k;
      atomic_add(t.gen[idx], d_tile * base);
      n++;
    } else {
      return n;
    }
    remaining -= d_tile;
    if (remaining <= 0.) break;
    r = add(r, scale(u, d_tile));
    on[0] = on[1] = on[2] = 0;
    switch (nt.second) {
      case -1:
        i--;
        on[0] = 1;
        break;
      case 1:
        i++;
        on[0] = -1;
        break;
      case -2:
        j--;
        on[1] = 1;
        break;
      case 2:
        j++;
        on[1] = -1;
        break;
      case -3:
        k--;
        on[2] = 1;
        break;
      case 3:
        k++;
        on[2] = -1;
        break;
      default:
        break;
    }
  }
  return n;
}

//==========================================================================
// Per-history context
struct Ctx {
  Problem* P;
  double kcol;
  bool converged;
  Scores sc;
  abx_diag diag{};
  std::vector<abx_trace_rec>* trace = nullptr;
  int trace_cap = 0;
  int32_t n_rec = 0;
  uint64_t rng0_calls = 0;

  void rec(const Particle& p, const Tracker& t, int ev) {
    if (!trace) return;
    if (n_rec < trace_cap) {
      abx_trace_rec r{};
      r.r[0] = p.s.r.x;
      r.r[1] = p.s.r.y;
      r.r[2] = p.s.r.z;
      r.u[0] = p.s.u.x;
      r.u[1] = p.s.u.y;
      r.u[2] = p.s.u.z;
      r.wgt = p.s.w;
      r.history_id = p.history_id;
      r.secondary_id = (int32_t)p.secondary_id;
      r.event = ev;
      r.cell = t.cell;
      r.material = t.material();
      r.group = energy_grid_index(*P, p.s.E);
      r.rng_calls = (int32_t)(p.rng.calls - rng0_calls);
      trace->push_back(r);
    }
    n_rec++;
  }
};

static void tally_collision_all(Ctx& c, const Particle& p, int mi) {
  if (!c.converged) return;
  for (auto& t : c.P->tallies)
    if (t.estimator == ABX_EST_COLLISION)
      c.diag.tally_updates += score_collision(*c.P, t, p, mi);
}

static void tally_flight_all(Ctx& c, const Particle& p, double d, int mi) {
  if (!c.converged) return;
  for (auto& t : c.P->tallies)
    if (t.estimator == ABX_EST_TRACK_LENGTH)
      c.diag.tally_updates += score_flight(*c.P, t, p, d, mi);
}

//==========================================================================
// Transporter physics (src/transporter.cpp)
static void russian_roulette(Ctx& c, Particle& p) {  // :35-57
  const abx_settings& st = c.P->st;
  if (std::abs(p.s.w) < st.wgt_cutoff) {
    double P_kill = 1.0 - (std::abs(p.s.w) / st.wgt_survival);
    if (RNG::rand(p.rng) < P_kill)
      p.s.w = 0.;
    else
      p.s.w = std::copysign(st.wgt_survival, p.s.w);
  }
  if (std::abs(p.s.w2) < st.wgt_cutoff) {
    double P_kill = 1.0 - (std::abs(p.s.w2) / st.wgt_survival);
    if (RNG::rand(p.rng) < P_kill)
      p.s.w2 = 0.;
    else
      p.s.w2 = std::copysign(st.wgt_survival, p.s.w2);
  }
  if (p.s.w == 0. && p.s.w2 == 0.) p.alive = false;
}

static void make_fission_neutrons(Ctx& c, Particle& p, const MicroXS& x,
                                  const Material& m) {  // :358-487
  const abx_settings& st = c.P->st;
  double k_abs_scr = p.s.w * x.nu_total * x.fission / x.total;
  int n_new = 0;
  if (st.mode == ABX_K_EIGENVALUE)
    n_new = (int)std::floor(std::abs(k_abs_scr) / c.kcol + RNG::rand(p.rng));
  else
    n_new = (int)std::floor(std::abs(k_abs_scr) + RNG::rand(p.rng));
  double P_delayed = x.nu_delayed / x.nu_total;
  for (int i = 0; i < n_new; i++) {
    // MGNuclide::sample_fission (src/mg_nuclide.cpp:504-544)
    const auto& chi = m.chi[(size_t)x.energy_index];
    size_t ei = (size_t)RNG::discrete(p.rng, chi.data(), chi.data() + chi.size());
    double E_out = 0.5 * (c.P->ebounds[ei] + c.P->ebounds[ei + 1]);
    double mu = 2. * RNG::rand(p.rng) - 1.;
    double phi = 2. * PI * RNG::rand(p.rng);
    V3 uout = rotate_direction(p.s.u, mu, phi);
    if (RNG::rand(p.rng) < P_delayed) {
      (void)RNG::discrete(p.rng, m.P_delayed_group.data(),
                          m.P_delayed_group.data() + m.P_delayed_group.size());
    }
    double wgt = p.s.w > 0. ? 1. : -1.;
    double wgt2 = 0.;
    if (st.mode == ABX_FIXED_SOURCE) {
      p.make_secondary(uout, E_out, wgt, wgt2);
    } else {
      abx_banked b{};
      b.r[0] = p.s.r.x;
      b.r[1] = p.s.r.y;
      b.r[2] = p.s.r.z;
      b.u[0] = uout.x;
      b.u[1] = uout.y;
      b.u[2] = uout.z;
      b.E = E_out;
      b.wgt = wgt;
      b.wgt2 = wgt2;
      b.parent_history_id = p.history_id;
      b.parent_daughter_id = p.daughter_counter++;
      b.family_id = p.family_id;
      p.fission.push_back(b);
      // the cancellation fields of the banked particle (:440-456)
      abx_parent_info q{};
      q.r1[0] = p.prev_r.x;
      q.r1[1] = p.prev_r.y;
      q.r1[2] = p.prev_r.z;
      q.u1[0] = p.prev_u.x;
      q.u1[1] = p.prev_u.y;
      q.u1[2] = p.prev_u.z;
      q.E1 = p.prev_E;
      q.E3 = p.s.E;
      q.Esmp = p.Esmp;
      q.prev_virtual = p.prev_virtual ? 1 : 0;
      p.fission_parent.push_back(q);
    }
  }
}

static void do_scatter(Ctx& c, Particle& p, const Material& m,
                       const MicroXS& x) {  // :317-347, mg_nuclide.cpp:442-461
  const auto& ps = m.Ps[(size_t)x.energy_index];
  size_t ei = (size_t)RNG::discrete(p.rng, ps.data(), ps.data() + ps.size());
  double E_out = 0.5 * (c.P->ebounds[ei] + c.P->ebounds[ei + 1]);
  double mu = m.angle[(size_t)x.energy_index][ei].sample_mu(p.rng);
  double phi = 2. * PI * RNG::rand(p.rng);
  V3 u_out = rotate_direction(p.s.u, mu, phi);
  p.set_direction(u_out);
  p.set_energy(E_out);
  p.s.w = p.s.w * 1.;
  p.s.w2 = p.s.w2 * 1.;
}

// Transporter::collision + branching_collision (:60-93, :269-315)
static void collision(Ctx& c, Particle& p, int mi) {
  const Problem& P = *c.P;
  tally_collision_all(c, p, mi);
  {
    double k_col_scr = p.s.w * mat_vEf(P, mi, p.s.E) / mat_Et(P, mi, p.s.E);
    double mig_dist = norm(sub(p.s.r, p.r_birth));
    double mig_area_scr =
        p.s.w * mat_Ea(P, mi, p.s.E) / mat_Et(P, mi, p.s.E) * mig_dist * mig_dist;
    c.sc.k_col += k_col_scr;
    c.sc.mig += mig_area_scr;
  }
  const Material& m = P.materials[(size_t)mi];
  // MaterialHelper::sample_nuclide: one draw, single component
  (void)RNG::rand(p.rng);
  MicroXS x = micro_xs(P, m, p.s.E);
  c.sc.k_abs += p.s.w * x.nu_total * x.fission / x.total;
  make_fission_neutrons(c, p, x, m);
  p.s.w = p.s.w * (1. - (x.absorption + 0.) / x.total);
  p.s.w2 = p.s.w2 * (1. - (x.absorption + 0.) / x.total);
  russian_roulette(c, p);
  if (p.alive) {
    do_scatter(c, p, m, x);
    if (p.s.E < P.min_energy) p.alive = false;
  }
}

static void leak_score(Ctx& c, Particle& p, const BoundaryInfo& b) {
  p.alive = false;
  c.sc.leak += p.s.w;
  V3 r_leak = add(p.s.r, scale(p.s.u, b.distance));
  V3 dr = sub(r_leak, p.r_birth);
  // w * (r_leak - r_birth) * (r_leak - r_birth): scale first, then dot
  c.sc.mig += dot(scale(dr, p.s.w), dr);
}

// Shared tail of each loop iteration: resurrection of secondaries.
static void try_resurrect(Ctx& c, Particle& p, Tracker& t, int& mi) {
  if (p.alive) return;
  p.resurect();
  if (p.alive) {
    c.diag.secondaries++;
    t.set_r(p.s.r);
    t.set_u(p.s.u);
    t.restart_get_current();
    if (t.is_lost()) fatal(ABX_E_LOST, "lost at resurrection");
    mi = t.material();
    c.rec(p, t, ABX_EV_RESURRECT);
  }
}

// DeltaTracker::transport loop body (src/delta_tracker.cpp:86-229) and
// CarterTracker (src/carter_tracker.cpp:104-260)
// Flights per history (with its resurrected secondaries) before it is
// reported lost instead of looping for ever; the library's kernels stop at
// the same count.
constexpr long kMaxFlights = 100000000;

static void run_history_dt(Ctx& c, Particle& p, bool carter) {
  const Problem& P = *c.P;
  Tracker t(&P, p.s.r, p.s.u);
  if (t.is_lost()) {
    p.alive = false;
    c.diag.lost++;
  }
  int mi = t.material();
  long steps = 0;
  while (p.alive) {
    if (++steps > kMaxFlights) fatal(ABX_E_LOST, "history exceeds the flight limit");
    bool had_collision = false;
    bool crossed = false;
    int g = energy_grid_index(P, p.s.E);
    double Emaj = (carter ? P.sampling[(size_t)g] : P.majorant[(size_t)g]) + 0.;
    p.Esmp = Emaj;
    double d_coll = RNG::exponential(p.rng, Emaj);
    c.diag.flights++;
    BoundaryInfo bound{INF, -1, ABX_NORMAL, 0};
    t.move(d_coll);
    t.get_current();
    if (t.is_lost()) {
      t.set_r(p.s.r);
      t.get_current();
      bound = t.get_boundary_condition();
      crossed = true;
    }
    tally_flight_all(c, p, std::min(d_coll, bound.distance), mi);
    int ev = ABX_EV_VIRTUAL;
    if (crossed) {
      c.diag.crossings++;
      if (bound.btype == ABX_VACUUM) {
        leak_score(c, p, bound);
        ev = ABX_EV_LEAK;
      } else if (bound.btype == ABX_REFLECTIVE) {
        t.do_reflection(p, bound);
        if (t.is_lost()) fatal(ABX_E_LOST, "lost after reflection");
        ev = ABX_EV_REFLECT;
      } else {
        fatal(ABX_E_LOST, "Help me, how did I get here ?");
      }
    } else {
      p.move(d_coll);
      mi = t.material();
      double Et = mat_Et(P, mi, p.s.E);
      if (!carter) {
        if (Et - Emaj > 1.E-10) fatal(ABX_E_MAJORANT, "Sigma_t exceeded majorant");
        if (RNG::rand(p.rng) < (Et / Emaj)) had_collision = true;
      } else {
        if (Emaj >= Et) {
          if (RNG::rand(p.rng) < (Et / Emaj)) had_collision = true;
        } else {
          double D = Et / (2. * Et - Emaj);
          double F = Et / (D * Emaj);
          if ((D - RNG::rand(p.rng)) > 0.) {
            p.s.w = p.s.w * F;
            had_collision = true;
          } else {
            p.s.w = -p.s.w * F;
          }
        }
      }
    }
    if (p.alive && had_collision) {
      c.diag.collisions++;
      collision(c, p, mi);
      t.set_u(p.s.u);
      p.prev_virtual = false;
      ev = p.alive ? ABX_EV_COLLISION : ABX_EV_DEATH;
    } else if (p.alive) {
      if (!crossed) c.diag.virtual_collisions++;
      p.prev_virtual = true;
    }
    if (carter && p.alive && std::abs(p.s.w) >= P.st.wgt_split) {
      int n_new = (int)std::ceil(std::abs(p.s.w));
      p.split(n_new);
    }
    c.rec(p, t, ev);
    try_resurrect(c, p, t, mi);
  }
}

// SurfaceTracker::transport loop body (src/surface_tracker.cpp:70-200)
static void run_history_st(Ctx& c, Particle& p) {
  const Problem& P = *c.P;
  Tracker t(&P, p.s.r, p.s.u);
  if (t.is_lost()) {
    p.alive = false;
    c.diag.lost++;
  }
  int mi = t.material();
  long steps = 0;
  while (p.alive) {
    if (++steps > kMaxFlights) fatal(ABX_E_LOST, "history exceeds the flight limit");
    bool had_collision = false;
    double d_coll = RNG::exponential(p.rng, mat_Et(P, mi, p.s.E));
    c.diag.flights++;
    BoundaryInfo bound = t.get_nearest_boundary();
    double dmin = std::min(d_coll, bound.distance);
    tally_flight_all(c, p, dmin, mi);
    c.sc.k_trk += p.s.w * dmin * mat_vEf(P, mi, p.s.E);
    int ev = ABX_EV_CROSS;
    if (bound.distance < d_coll || std::abs(bound.distance - d_coll) < BOUNDRY_TOL) {
      c.diag.crossings++;
      if (bound.btype == ABX_VACUUM) {
        leak_score(c, p, bound);
        ev = ABX_EV_LEAK;
      } else if (bound.btype == ABX_REFLECTIVE) {
        t.do_reflection(p, bound);
        if (t.is_lost()) fatal(ABX_E_LOST, "lost after reflection");
        ev = ABX_EV_REFLECT;
      } else {
        t.cross_surface(bound);
        t.get_current();
        p.move(bound.distance);
        if (t.is_lost()) fatal(ABX_E_LOST, "lost after crossing");
        mi = t.material();
      }
    } else {
      p.move(d_coll);
      t.move(d_coll);
      had_collision = true;
    }
    if (p.alive && had_collision) {
      c.diag.collisions++;
      collision(c, p, mi);
      t.set_u(p.s.u);
      ev = p.alive ? ABX_EV_COLLISION : ABX_EV_DEATH;
    }
    c.rec(p, t, ev);
    try_resurrect(c, p, t, mi);
  }
}

//==========================================================================
// Source sampling (src/simulation.cpp:55-77, src/source.cpp:44-90)
static abx_particle sample_one(Problem& P, uint64_t h) {
  Pcg32 rng;
  rng.seed(P.st.seed);
  rng.advance(P.st.stride * h);
  std::vector<double> wgts;
  for (auto& s : P.sources) wgts.push_back(s.d.weight);
  size_t indx = (size_t)RNG::discrete(rng, wgts.data(), wgts.data() + wgts.size());
  const abx_source& s = P.sources[indx].d;
  V3 u;
  if (s.direction == ABX_DIR_ISOTROPIC) {
    double mu = 2. * RNG::rand(rng) - 1.;
    double phi = 2. * PI * RNG::rand(rng);
    u = direction_mu_phi(mu, phi);
  } else {
    u = direction(s.dir[0], s.dir[1], s.dir[2]);
  }
  double E = s.energy;
  int E_count = 0;
  do {
    if (E_count > 200) fatal(ABX_E_INVALID, "Exceded 200 samplings of energy.");
    E = s.energy;
    E_count++;
  } while (E <= P.min_energy || P.max_energy <= E);
  auto sample_r = [&](Pcg32& g) -> V3 {
    if (s.spatial == ABX_SRC_POINT) return {s.low[0], s.low[1], s.low[2]};
    double x = (s.high[0] - s.low[0]) * RNG::rand(g) + s.low[0];
    double y = (s.high[1] - s.low[1]) * RNG::rand(g) + s.low[1];
    double z = (s.high[2] - s.low[2]) * RNG::rand(g) + s.low[2];
    return {x, y, z};
  };
  V3 r = sample_r(rng);
  Tracker t(&P, r, u);
  int guard = 0;
  while (t.is_lost()) {
    if (++guard > 100000) fatal(ABX_E_LOST, "source never inside geometry");
    r = sample_r(rng);
    t.set_r(r);
    t.restart_get_current();
  }
  if (s.fissile_only) {
    int count = 0;
    while (t.is_lost() || !P.materials[(size_t)t.material()].fissile) {
      if (count == 201)
        fatal(ABX_E_INVALID, "Exceded 200 samplings of position in fissile-only source.");
      r = sample_r(rng);
      t.set_r(r);
      t.restart_get_current();
      count++;
    }
  }
  abx_particle out{};
  out.r[0] = r.x;
  out.r[1] = r.y;
  out.r[2] = r.z;
  out.u[0] = u.x;
  out.u[1] = u.y;
  out.u[2] = u.z;
  out.E = E;
  out.wgt = 1.0;
  out.wgt2 = 0.;
  out.history_id = h;
  out.family_id = h;
  out.rng_state = rng.state;
  return out;
}

//==========================================================================
// Problem construction (validation as MGNuclide checks, src/mg_nuclide.cpp:235-282)
static void build(Problem& P, const abx_problem_desc* d) {
  P.st = d->settings;
  const int G = d->settings.ngroups;
  if (G <= 0 || !d->settings.energy_bounds) fatal(ABX_E_INVALID, "ngroups");
  P.ebounds.assign(d->settings.energy_bounds, d->settings.energy_bounds + G + 1);
  for (int i = 0; i < d->n_surfaces; i++) {
    const abx_surface& s = d->surfaces[i];
    Surface S;
    S.type = s.type;
    S.bc = s.boundary;
    S.id = s.id;
    for (int k = 0; k < 4; k++) S.p[k] = s.p[k];
    if (s.type == ABX_CYL) {
      // Cylinder::Cylinder, src/cylinder.cpp:42-57
      S.x0 = s.p[0];
      S.y0 = s.p[1];
      S.z0 = s.p[2];
      double u0 = s.p[3], v0 = s.p[4], w0 = s.p[5];
      const double mag = std::sqrt(u0 * u0 + v0 * v0 + w0 * w0);
      if (mag == 0.)
        fatal(ABX_E_INVALID, "Direction vector of Cylinder can not have a mangitude of 0.");
      u0 /= mag;
      v0 /= mag;
      w0 /= mag;
      S.alpha = 1. - u0 * u0;
      S.beta = 1. - v0 * v0;
      S.gamma = 1. - w0 * w0;
      S.R = s.p[6];
    } else if (s.type < ABX_XPLANE || s.type > ABX_SPHERE) {
      fatal(ABX_E_INVALID, "surface type");
    }
    P.surfaces.push_back(S);
  }
  for (int i = 0; i < d->n_cells; i++) {
    const abx_cell& c = d->cells[i];
    Cell C;
    C.id = c.id;
    C.fill_type = c.fill_type;
    C.fill = c.fill;
    C.rpn.assign(d->rpn + c.rpn_offset, d->rpn + c.rpn_offset + c.rpn_len);
    // Cell::simplify (src/cell.cpp:210-232)
    C.simple = true;
    for (int32_t el : C.rpn)
      if (el == ABX_OP_COMP || el == ABX_OP_UNIN) {
        C.simple = false;
        break;
      }
    if (C.simple) {
      std::vector<int32_t> t;
      for (int32_t el : C.rpn)
        if (el < ABX_OP_UNIN) t.push_back(el);
      C.rpn = t;
    }
    // Cell::check_for_bc (src/cell.cpp:196-208)
    for (int32_t token : C.rpn) {
      if (token >= ABX_OP_UNIN) continue;
      int b = P.surfaces[(size_t)(std::abs(token) - 1)].bc;
      if (b == ABX_VACUUM || b == ABX_REFLECTIVE) C.vacuum_or_reflective = true;
    }
    P.cells.push_back(C);
  }
  for (int i = 0; i < d->n_universes; i++) {
    const abx_universe& u = d->universes[i];
    Universe U;
    U.id = u.id;
    U.type = u.type;
    if (u.type == ABX_CELL_UNIVERSE) {
      U.cells.assign(d->universe_cells + u.cells_offset,
                     d->universe_cells + u.cells_offset + u.n_cells);
    } else if (u.type == ABX_HEX_LATTICE) {
      // HexLattice constructor and set_elements (src/hex_lattice.cpp:30-56,
      // 219-243); the trigonometric constants as in
      // include/geometry/hex_lattice.hpp:60-63
      U.rings = (uint32_t)u.shape[0];
      U.Nz = (uint32_t)u.shape[1];
      U.Px = u.pitch[0];
      U.Pz = u.pitch[1];
      U.Xl = u.origin[0];
      U.Yl = u.origin[1];
      U.Zl = u.origin[2];
      U.top = u.hex_top;
      U.c6 = std::cos(PI / 6.0);
      U.s6 = std::sin(PI / 6.0);
      U.c3 = std::cos(PI / 3.0);
      U.s3 = std::sin(PI / 3.0);
      U.width = 2 * (U.rings - 1) + 1;
      U.mid = U.width / 2;
      U.Nx = U.Ny = U.width;
      U.elems.assign((size_t)U.width * U.width * U.Nz, -1);
      size_t k = 0;
      for (uint32_t az = 0; az < U.Nz; az++)
        for (uint32_t ar = 0; ar < U.width; ar++)
          for (uint32_t aq = 0; aq < U.width; aq++) {
            const int32_t q = (int32_t)(aq - U.mid), r = (int32_t)(ar - U.mid);
            if (hex_ring(q, r) < U.rings)
              U.elems[(size_t)az * U.width * U.width + ar * U.width + aq] =
                  d->lattice_elems[u.elems_offset + (int64_t)k++];
          }
      U.outer = u.outer;
    } else {
      U.Nx = (uint32_t)u.shape[0];
      U.Ny = (uint32_t)u.shape[1];
      U.Nz = (uint32_t)u.shape[2];
      U.Px = u.pitch[0];
      U.Py = u.pitch[1];
      U.Pz = u.pitch[2];
      U.Px_inv = 1. / U.Px;
      U.Py_inv = 1. / U.Py;
      U.Pz_inv = 1. / U.Pz;
      U.Xl = u.origin[0] - (double)U.Nx * 0.5 * U.Px;
      U.Yl = u.origin[1] - (double)U.Ny * 0.5 * U.Py;
      U.Zl = u.origin[2] - (double)U.Nz * 0.5 * U.Pz;
      size_t ne = (size_t)U.Nx * U.Ny * U.Nz;
      U.elems.assign(d->lattice_elems + u.elems_offset,
                     d->lattice_elems + u.elems_offset + ne);
      U.outer = u.outer;
    }
    P.universes.push_back(U);
  }
  // has_boundary_conditions: cell universes from their cells
  // (src/cell_universe.cpp:30-41), lattices from the outer universe
  // (src/lattice.cpp:35-52) - lattices are built after the universes they
  // reference, so resolve cell universes first.
  for (auto& U : P.universes)
    if (U.type == ABX_CELL_UNIVERSE)
      for (int ci : U.cells)
        if (P.cells[(size_t)ci].vacuum_or_reflective) U.has_bc = true;
  for (int pass = 0; pass < 8; pass++)
    for (auto& U : P.universes)
      if (U.type != ABX_CELL_UNIVERSE && U.outer >= 0)
        U.has_bc = P.universes[(size_t)U.outer].has_bc;
  P.root = d->root_universe;

  for (int mi = 0; mi < d->n_materials; mi++) {
    const abx_material& a = d->materials[mi];
    Material M;
    M.id = a.id;
    M.G = G;
    M.Et.assign(a.total, a.total + G);
    M.Ea.assign(a.absorption, a.absorption + G);
    M.Ef.assign((size_t)G, 0.);
    M.nu_p.assign((size_t)G, 0.);
    M.nu_d.assign((size_t)G, 0.);
    bool fis = false;
    if (a.fission) {
      M.Ef.assign(a.fission, a.fission + G);
      for (double f : M.Ef)
        if (f > 0.) fis = true;
    }
    if (fis) {
      if (!a.nu_prompt) fatal(ABX_E_INVALID, "No nu data");
      M.nu_p.assign(a.nu_prompt, a.nu_prompt + G);
      if (a.nu_delayed) {
        M.nu_d.assign(a.nu_delayed, a.nu_delayed + G);
      }
    }
    // nu_delyd_ is always sized G (zeros without delayed data), so nu_total
    // adds it (src/mg_nuclide.cpp:353-359)
    M.has_nu_d = true;
    M.chi.assign((size_t)G, std::vector<double>((size_t)G, 0.));
    if (fis) {
      if (!a.chi) fatal(ABX_E_INVALID, "Invalid chi matrix");
      for (int i = 0; i < G; i++)
        for (int o = 0; o < G; o++) M.chi[(size_t)i][(size_t)o] = a.chi[i * G + o];
    }
    // MGNuclide::normalize_chi (:73-93)
    for (int i = 0; i < G; i++) {
      double chi_i = 0.;
      for (int o = 0; o < G; o++) chi_i += M.chi[(size_t)i][(size_t)o];
      for (int o = 0; o < G; o++) M.chi[(size_t)i][(size_t)o] /= chi_i;
    }
    // make_scatter_xs (:95-118)
    M.Ps.assign((size_t)G, std::vector<double>((size_t)G, 0.));
    M.Es.assign((size_t)G, 0.);
    for (int i = 0; i < G; i++)
      for (int o = 0; o < G; o++) M.Ps[(size_t)i][(size_t)o] = a.scatter[i * G + o];
    for (int i = 0; i < G; i++) {
      M.Es[(size_t)i] = 0.;
      for (int o = 0; o < G; o++) M.Es[(size_t)i] += M.Ps[(size_t)i][(size_t)o];
      for (int o = 0; o < G; o++) M.Ps[(size_t)i][(size_t)o] /= M.Es[(size_t)i];
    }
    M.angle.assign((size_t)G, std::vector<AngleDist>((size_t)G));
    if (a.angle_offsets) {
      for (int i = 0; i < G; i++)
        for (int o = 0; o < G; o++) {
          int b0 = a.angle_offsets[i * G + o], b1 = a.angle_offsets[i * G + o + 1];
          AngleDist& ad = M.angle[(size_t)i][(size_t)o];
          ad.mu.assign(a.angle_mu + b0, a.angle_mu + b1);
          ad.pdf.assign(a.angle_pdf + b0, a.angle_pdf + b1);
          ad.cdf.assign(a.angle_cdf + b0, a.angle_cdf + b1);
        }
    }
    if (a.n_delayed > 0) {
      M.P_delayed_group.assign(a.delayed_probabilities,
                               a.delayed_probabilities + a.n_delayed);
      M.delayed_constants.assign(a.delayed_constants,
                                 a.delayed_constants + a.n_delayed);
    }
    // check_xs (:235-282)
    for (int i = 0; i < G; i++) {
      if (M.Ea[(size_t)i] < 0. || M.Ef[(size_t)i] < 0.)
        fatal(ABX_E_INVALID, "negative cross section");
      if (M.Ef[(size_t)i] > M.Ea[(size_t)i]) fatal(ABX_E_INVALID, "Ef > Ea");
      double diff = M.Et[(size_t)i] - (M.Es[(size_t)i] + M.Ea[(size_t)i]);
      if (std::abs(diff) / M.Et[(size_t)i] > 0.001)
        fatal(ABX_E_INVALID, "Es + Ea != Et");
    }
    // check_fissile (:323-330)
    for (int i = 0; i < G; i++)
      if (M.Ef[(size_t)i] != 0. && (M.nu_p[(size_t)i] + M.nu_d[(size_t)i]) > 0.) {
        M.fissile = true;
        break;
      }
    P.materials.push_back(M);
  }
  // make_materials: energy limits (src/parser.cpp:152-167)
  P.min_energy = std::max(0., P.ebounds.front());
  P.max_energy = std::min(100000., P.ebounds.back());
  // make_majorant_xs, MG branch (src/majorant.cpp:133-176)
  P.majorant.assign((size_t)G, 0.);
  for (const auto& M : P.materials)
    for (int g = 0; g < G; g++)
      if (M.Et[(size_t)g] > P.majorant[(size_t)g]) P.majorant[(size_t)g] = M.Et[(size_t)g];
  // CarterTracker sampling xs (src/carter_tracker.cpp:60-75)
  P.sampling = P.majorant;
  if (d->settings.sample_xs_ratio) {
    P.ratio.assign(d->settings.sample_xs_ratio, d->settings.sample_xs_ratio + G);
    for (int g = 0; g < G; g++) P.sampling[(size_t)g] = P.majorant[(size_t)g] * P.ratio[(size_t)g];
  }
  for (int i = 0; i < d->n_sources; i++) P.sources.push_back({d->sources[i]});
  if (d->cancel) {
    const abx_cancel_mesh& c = *d->cancel;
    P.cancel = true;
    P.c_low = {c.low[0], c.low[1], c.low[2]};
    P.c_hi = {c.high[0], c.high[1], c.high[2]};
    if (P.c_low.x >= P.c_hi.x || P.c_low.y >= P.c_hi.y || P.c_low.z >= P.c_hi.z)
      fatal(ABX_E_INVALID, "Low position is not lower than hi position in ApproximateMeshCancelator.");
    for (int k = 0; k < 3; k++) P.c_shape[k] = (uint32_t)c.shape[k];
    P.c_dx = (P.c_hi.x - P.c_low.x) / (double)P.c_shape[0];
    P.c_dy = (P.c_hi.y - P.c_low.y) / (double)P.c_shape[1];
    P.c_dz = (P.c_hi.z - P.c_low.z) / (double)P.c_shape[2];
    if (c.n_ebounds > 0) {
      if (c.n_ebounds < 2) fatal(ABX_E_INVALID, "energy_edges must have at least two entries");
      P.c_edges.assign(c.ebounds, c.ebounds + c.n_ebounds);
      P.c_shape[3] = (uint32_t)(P.c_edges.size() - 1);
    }
  }
  if (d->exact_cancel) {
    // make_cancelator (src/cancelator.cpp:32-78) and the cancelators'
    // constructors (src/exact_mg_cancelator.cpp:41-118,
    // src/basic_exact_mg_cancelator.cpp:43-66)
    const abx_exact_cancel& e = *d->exact_cancel;
    if (e.kind != ABX_EXACT_MG && e.kind != ABX_BASIC_EXACT_MG)
      fatal(ABX_E_INVALID, "Unkown cancelator type.");
    if (d->cancel) fatal(ABX_E_INVALID, "one cancelator per problem");
    if (P.st.tracking == ABX_SURFACE_TRACKING)
      fatal(ABX_E_INVALID, "exect cancelators may not be used with surface-tracking.");
    P.exact = e.kind;
    P.ex = e;
    P.ex_low = {e.low[0], e.low[1], e.low[2]};
    P.ex_hi = {e.high[0], e.high[1], e.high[2]};
    if (P.ex_low.x >= P.ex_hi.x || P.ex_low.y >= P.ex_hi.y || P.ex_low.z >= P.ex_hi.z)
      fatal(ABX_E_INVALID, "r_low >= r_hi");
    if (e.n_samples <= 0) fatal(ABX_E_INVALID, "n-samples must be greater than zero.");
    for (int k = 0; k < 3; k++) {
      if (e.shape[k] < 1) fatal(ABX_E_INVALID, "cancelator shape");
      P.ex_shape[k] = (uint32_t)e.shape[k];
    }
    P.ex_pitch[0] = (P.ex_hi.x - P.ex_low.x) / (double)P.ex_shape[0];
    P.ex_pitch[1] = (P.ex_hi.y - P.ex_low.y) / (double)P.ex_shape[1];
    P.ex_pitch[2] = (P.ex_hi.z - P.ex_low.z) / (double)P.ex_shape[2];
    if (e.kind == ABX_EXACT_MG) {
      std::vector<int> seen((size_t)G, 0);
      for (int b = 0; b < e.n_group_bins; b++) {
        std::vector<int> grp(e.group_bins + e.group_bin_offsets[b],
                             e.group_bins + e.group_bin_offsets[b + 1]);
        for (int g : grp) {
          if (g < 0 || g >= G) fatal(ABX_E_INVALID, "grp >= settings::ngroups");
          if (seen[(size_t)g]++) fatal(ABX_E_INVALID, "Group is included more than once in group_bins.");
        }
        // a uniform particle of an empty energy region would index no group
        if (grp.empty() && e.chi_matrix) fatal(ABX_E_INVALID, "empty group bin");
        P.ex_groups.push_back(grp);
      }
      if (e.chi_matrix && P.ex_groups.empty())
        fatal(ABX_E_INVALID, "Chi matrix is used, but no group_bins provided.");
      P.ex_shape[3] = P.ex_groups.empty() ? 1u : (uint32_t)P.ex_groups.size();
    } else if (e.beta < ABX_BETA_ZERO || e.beta > ABX_BETA_AVERAGE_G) {
      fatal(ABX_E_INVALID, "Unkown beta entry for basic exact MG cancelator.");
    }
  }
  for (int i = 0; i < d->n_tallies; i++) {
    const abx_mesh_tally& a = d->tallies[i];
    MeshTally T;
    T.estimator = a.estimator;
    T.quantity = a.quantity;
    T.Nx = a.shape[0];
    T.Ny = a.shape[1];
    T.Nz = a.shape[2];
    T.r_low = {a.low[0], a.low[1], a.low[2]};
    T.r_hi = {a.high[0], a.high[1], a.high[2]};
    // MeshTally ctor (src/mesh_tally.cpp:62-116)
    T.dx = (T.r_hi.x - T.r_low.x) / (double)T.Nx;
    T.dy = (T.r_hi.y - T.r_low.y) / (double)T.Ny;
    T.dz = (T.r_hi.z - T.r_low.z) / (double)T.Nz;
    T.dx_inv = 1. / T.dx;
    T.dy_inv = 1. / T.dy;
    T.dz_inv = 1. / T.dz;
    T.ebounds.assign(a.ebounds, a.ebounds + a.n_ebounds);
    size_t nb = (size_t)(a.n_ebounds - 1) * T.Nx * T.Ny * T.Nz;
    T.gen.assign(nb, 0.);
    T.avg.assign(nb, 0.);
    T.var.assign(nb, 0.);
    // Tallies::add_*_mesh_tally: net weight = total weight = nparticles
    T.net_weight = (double)d->settings.nparticles;
    P.tallies.push_back(T);
  }
}

static Particle make_particle(const abx_particle& a) {
  Particle p;
  p.s.r = {a.r[0], a.r[1], a.r[2]};
  p.s.u = {a.u[0], a.u[1], a.u[2]};
  p.s.E = a.E;
  p.s.w = a.wgt;
  p.s.w2 = a.wgt2;
  p.history_id = a.history_id;
  p.family_id = a.family_id;
  p.rng.state = a.rng_state;
  p.r_birth = p.s.r;
  return p;
}

static void transport(Problem& P, const abx_particle* bank, size_t n,
                      double kcol, bool converged, int nthreads,
                      std::vector<abx_banked>& fission_out, Scores& total,
                      abx_diag& diag, abx_trace_rec* trace, int cap,
                      int32_t* n_rec) {
  std::vector<std::vector<abx_banked>> fis(n);
  std::vector<std::vector<abx_parent_info>> par(n);
  std::string err;
  int errcode = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    Ctx c;
    c.P = &P;
    c.kcol = kcol;
    c.converged = converged;
    std::vector<abx_trace_rec> tr;
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
    for (size_t i = 0; i < n; i++) {
      if (errcode) continue;
      try {
        Particle p = make_particle(bank[i]);
        c.rng0_calls = p.rng.calls;
        if (trace) {
          tr.clear();
          c.trace = &tr;
          c.trace_cap = cap;
          c.n_rec = 0;
        }
        if (P.st.tracking == ABX_SURFACE_TRACKING)
          run_history_st(c, p);
        else
          run_history_dt(c, p, P.st.tracking == ABX_CARTER_TRACKING);
        c.diag.histories++;
        fis[i] = std::move(p.fission);
        par[i] = std::move(p.fission_parent);
        if (trace) {
          std::copy(tr.begin(), tr.end(), trace + i * (size_t)cap);
          n_rec[i] = c.n_rec;
        }
      } catch (const Fatal& f) {
#ifdef _OPENMP
#pragma omp critical
#endif
        {
          if (!errcode) {
            errcode = f.code;
            err = f.msg;
            diag.error_history = bank[i].history_id;
          }
        }
      }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    {
      total.k_col += c.sc.k_col;
      total.k_abs += c.sc.k_abs;
      total.k_trk += c.sc.k_trk;
      total.k_tot += c.sc.k_tot;
      total.leak += c.sc.leak;
      total.mig += c.sc.mig;
      diag.histories += c.diag.histories;
      diag.collisions += c.diag.collisions;
      diag.virtual_collisions += c.diag.virtual_collisions;
      diag.flights += c.diag.flights;
      diag.crossings += c.diag.crossings;
      diag.tally_updates += c.diag.tally_updates;
      diag.lost += c.diag.lost;
      diag.secondaries += c.diag.secondaries;
    }
  }
  if (errcode) {
    diag.error = errcode;
    fatal(errcode, err);
  }
  for (auto& v : fis) fission_out.insert(fission_out.end(), v.begin(), v.end());
  P.last_parents.clear();
  for (auto& v : par) P.last_parents.insert(P.last_parents.end(), v.begin(), v.end());
}

static void tally_clear(Problem& P) {
  P.gen_scores = Scores();
  for (auto& t : P.tallies) std::fill(t.gen.begin(), t.gen.end(), 0.);
}

// Tallies::update_avg_and_var (src/tallies.cpp:268-280)
static void upd(int gen, double x, double& avg, double& var) {
  double dgen = (double)gen;
  double avg_old = avg, var_old = var;
  avg = avg_old + (x - avg_old) / dgen;
  if (gen > 1) var = var_old + ((x - avg_old) * (x - avg_old) / dgen) - (var_old / (dgen - 1.));
}

// Tallies::calc_gen_values (src/tallies.cpp:159-181): the generation's
// values join the per-generation vectors, recorded or not.
static void push_gen_values(Problem& P, const Scores& sc, double W) {
  const double x[6] = {sc.k_col / W, sc.k_abs / W, sc.k_trk / W,
                       sc.k_tot / W, sc.leak / W,  sc.mig / W};
  for (int q = 0; q < 6; q++) P.gen_vals[q].push_back(x[q]);
}

// Tallies::record_generation (src/tallies.cpp:183-203): the six scalar
// scores of the generation (already divided by the total weight, as
// calc_gen_values does, :159-181), then every mesh tally.
static void record_generation(Problem& P, const Scores& sc, double W) {
  P.gen++;
  upd(P.gen, sc.k_col / W, P.k_col_avg, P.k_col_var);
  upd(P.gen, sc.k_abs / W, P.k_abs_avg, P.k_abs_var);
  upd(P.gen, sc.k_trk / W, P.k_trk_avg, P.k_trk_var);
  upd(P.gen, sc.leak / W, P.leak_avg, P.leak_var);
  upd(P.gen, sc.k_tot / W, P.k_tot_avg, P.k_tot_var);
  upd(P.gen, sc.mig / W, P.mig_avg, P.mig_var);
  for (auto& t : P.tallies) {  // MeshTally::record_generation (:121-150)
    t.g++;
    const double dg = (double)t.g;
    for (size_t i = 0; i < t.gen.size(); i++) {
      double old_avg = t.avg[i];
      double val = t.gen[i] * 1.0;
      double avg = old_avg + (val - old_avg) / dg;
      t.avg[i] = avg;
      double var = t.var[i];
      var = var + (((val - old_avg) * (val - avg) - (var)) / dg);
      t.var[i] = var;
    }
  }
}

static void add_diag(abx_diag& a, const abx_diag& b) {
  a.histories += b.histories;
  a.collisions += b.collisions;
  a.virtual_collisions += b.virtual_collisions;
  a.flights += b.flights;
  a.crossings += b.crossings;
  a.tally_updates += b.tally_updates;
  a.lost += b.lost;
  a.secondaries += b.secondaries;
}

//==========================================================================
// Exact MG cancellation: ExactMGCancelator (src/exact_mg_cancelator.cpp)
// and BasicExactMGCancelator (src/basic_exact_mg_cancelator.cpp).
//
// The reference keeps its regions in nested std::unordered_maps keyed by the
// region key (i, j, k, e) and by Material*, so the order in which regions
// draw from settings::rng depends on the library's hashing and on heap
// addresses.  The restatement visits them in a fixed order - (i, j, k, e)
// lexicographically, then material index - and the library's device
// implementation visits them in the same order; everything else follows the
// reference step by step.

// sobol::sample (vendor/sobol/include/sobol/sobol.hpp): 52-bit generator
// matrices of the first four dimensions of the Joe-Kuo sequence
// (new-joe-kuo-6.21201), built from its primitive polynomials and initial
// direction numbers: dimension 0 is the van der Corput sequence, then
// (s, a, m) = (1, 0, {1}), (2, 1, {1, 3}), (3, 1, {1, 3, 1}).
static const uint64_t* sobol_matrices() {
  static uint64_t M[4 * 52];
  static bool built = false;
  if (built) return M;
  for (int k = 0; k < 52; k++) M[k] = 1ull << (51 - k);
  const int S[3] = {1, 2, 3}, A[3] = {0, 1, 1};
  const int Mi[3][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
  for (int d = 0; d < 3; d++) {
    uint64_t* V = M + (d + 1) * 52;
    const int s = S[d], a = A[d];
    for (int k = 0; k < s; k++) V[k] = (uint64_t)Mi[d][k] << (51 - k);
    for (int k = s; k < 52; k++) {
      uint64_t v = V[k - s] ^ (V[k - s] >> s);
      for (int j = 1; j < s; j++)
        if ((a >> (s - 1 - j)) & 1) v ^= V[k - j];
      V[k] = v;
    }
  }
  built = true;
  return M;
}

static double sobol_sample(uint64_t index, unsigned dim) {
  const uint64_t* M = sobol_matrices() + dim * 52;
  uint64_t r = 0;
  for (unsigned i = 0; index; index >>= 1, ++i)
    if (index & 1) r ^= M[i];
  return (double)r * (1.0 / (double)(1ull << 52));
}

// settings::group (src/settings.cpp:95-104): bounds inclusive at both ends
static size_t settings_group(const Problem& P, double E) {
  if (P.ebounds.size() <= 1) return 0;
  for (size_t g = 0; g < P.ebounds.size() - 1; g++)
    if (P.ebounds[g] <= E && E <= P.ebounds[g + 1]) return g;
  return 0;
}

// MGAngleDistribution::pdf (include/materials/mg_angle_distribution.hpp:62-75)
static double angle_pdf(const AngleDist& a, double mu) {
  if (mu < a.mu.front()) return a.pdf.front();
  if (mu > a.mu.back()) return a.pdf.back();
  auto it = std::lower_bound(a.mu.begin(), a.mu.end(), mu);
  size_t l = (size_t)std::distance(a.mu.begin(), it);
  if (mu == *it) return a.pdf[l];
  l--;
  const double m = (a.pdf[l + 1] - a.pdf[l]) / (a.mu[l + 1] - a.mu[l]);
  return m * (mu - a.mu[l]) + a.pdf[l];
}

// get_material (src/exact_mg_cancelator.cpp:154-170): the material index at
// r, located with u = (1, 0, 0); -1 outside every cell
static int material_at(const Problem& P, const V3& r) {
  Tracker t(&P, r, {1., 0., 0.});
  if (t.cell < 0) return -1;
  return t.material();
}

struct ExBin {  // CancelBin (include/simulation/exact_mg_cancelator.hpp:54-68)
  double uniform_wgt = 0., uniform_wgt2 = 0., W = 0., W2 = 0.;
  double sum_c = 0., sum_c_wgt = 0., sum_c_wgt2 = 0.;
  uint64_t rng_seed_advance = 0;
  bool can_cancel = true;
  std::vector<size_t> particles;  // bank indices, in bank order
  std::vector<double> f, f_inv;   // Averages
};
struct ExKey {
  size_t i, j, k, e;
};
// keyed like the reference's maps: Key equality on (i, j, k, e), then the
// material (a site on the mesh's upper faces has i, j or k equal to the
// shape, and a group in no group bin has e equal to their number)
using ExBins =
    std::map<std::tuple<size_t, size_t, size_t, size_t, int>, std::pair<ExKey, ExBin>>;

// ExactMGCancelator::get_f (:220-241)
static double exact_get_f(const Problem& P, const Material& m, const V3& r1, const V3& u1,
                          size_t g1, size_t g3, const V3& r4, size_t g4, double Esmp) {
  const double d = norm(sub(r4, r1));
  const V3 u = direction(r4.x - r1.x, r4.y - r1.y, r4.z - r1.z);
  const double mu = dot(u, u1);
  const double pdf_mu = angle_pdf(m.angle[g1][g3], mu);
  const double pdf_chi = P.ex.chi_matrix ? m.chi[g3][g4] : 1.;
  return (pdf_mu * pdf_chi / (d * d)) * EXP(-Esmp * d);
}

// BasicExactMGCancelator::get_f (:188-216)
static double basic_get_f(const V3& r, const V3& rp, double Esmp) {
  const double d = std::sqrt(std::pow(r.x - rp.x, 2.) + std::pow(r.y - rp.y, 2.) +
                             std::pow(r.z - rp.z, 2.));
  return (1. / (d * d)) * EXP(-Esmp * d);
}

static V3 v3(const double* a) { return {a[0], a[1], a[2]}; }

// ExactMGCancelator::get_key (:121-152); nullopt -> false
static bool exact_get_key(const Problem& P, const V3& r, size_t g, ExKey& key) {
  if (r.x < P.ex_low.x || r.x > P.ex_hi.x || r.y < P.ex_low.y || r.y > P.ex_hi.y ||
      r.z < P.ex_low.z || r.z > P.ex_hi.z)
    return false;
  key.i = (size_t)std::floor((r.x - P.ex_low.x) / P.ex_pitch[0]);
  key.j = (size_t)std::floor((r.y - P.ex_low.y) / P.ex_pitch[1]);
  key.k = (size_t)std::floor((r.z - P.ex_low.z) / P.ex_pitch[2]);
  bool found = false;
  size_t e = 0;
  for (e = 0; e < P.ex_groups.size(); e++) {
    for (size_t i = 0; i < P.ex_groups[e].size(); i++)
      if ((int)g == P.ex_groups[e][i]) {
        found = true;
        break;
      }
    if (found) break;
  }
  if (!found && P.ex.chi_matrix) return false;
  key.e = e;
  return true;
}

// Sample a point of a region in a material: sobol index `si` or the rng.
static bool region_point(const Problem& P, const ExKey& key, int mat, uint64_t* si, Pcg32* g,
                         V3& r) {
  const double Xl = P.ex_low.x + (double)key.i * P.ex_pitch[0];
  const double Yl = P.ex_low.y + (double)key.j * P.ex_pitch[1];
  const double Zl = P.ex_low.z + (double)key.k * P.ex_pitch[2];
  for (uint32_t tries = 0; tries < 100; tries++) {  // N_MAX_POS
    double x, y, z;
    if (si) {
      x = Xl + sobol_sample(*si, 0) * P.ex_pitch[0];
      y = Yl + sobol_sample(*si, 1) * P.ex_pitch[1];
      z = Zl + sobol_sample(*si, 2) * P.ex_pitch[2];
      (*si)++;
    } else {
      x = Xl + RNG::rand(*g) * P.ex_pitch[0];
      y = Yl + RNG::rand(*g) * P.ex_pitch[1];
      z = Zl + RNG::rand(*g) * P.ex_pitch[2];
    }
    r = {x, y, z};
    if (material_at(P, r) == mat) return true;
  }
  return false;
}

// Distribute the bank into regions (add_particle, exact :172-209 / basic
// :68-113).
static ExBins exact_bins(const Problem& P, const std::vector<abx_banked>& bank,
                         const std::vector<abx_parent_info>& par) {
  ExBins bins;
  for (size_t n = 0; n < bank.size(); n++) {
    const abx_banked& b = bank[n];
    const V3 r = v3(b.r);
    ExKey key{0, 0, 0, 0};
    if (P.exact == ABX_EXACT_MG) {
      if (!exact_get_key(P, r, settings_group(P, b.E), key)) continue;
      if (!P.ex.use_virtual_collisions && par[n].prev_virtual) continue;
    } else {
      const int i = (int)std::floor((r.x - P.ex_low.x) / P.ex_pitch[0]);
      const int j = (int)std::floor((r.y - P.ex_low.y) / P.ex_pitch[1]);
      const int k = (int)std::floor((r.z - P.ex_low.z) / P.ex_pitch[2]);
      if (i < 0 || j < 0 || k < 0) continue;
      if (i >= (int)P.ex_shape[0] || j >= (int)P.ex_shape[1] || k >= (int)P.ex_shape[2]) continue;
      key = {(size_t)i, (size_t)j, (size_t)k, 0};
    }
    // get_material (:154-170) has no material outside every cell; the bin
    // would then dereference none.  A fission site is always in a cell.
    const int mat = material_at(P, r);
    if (mat < 0) fatal(ABX_E_INVALID, "exact cancellation: fission site outside every cell");
    auto& kb = bins[{key.i, key.j, key.k, key.e, mat}];
    kb.first = key;
    kb.second.particles.push_back(n);
    kb.second.W += b.wgt;
    kb.second.W2 += b.wgt2;
  }
  return bins;
}

static void mixed_signs(const std::vector<abx_banked>& bank, const ExBin& bin, bool& m1,
                        bool& m2) {
  bool p1 = false, n1 = false, p2 = false, n2 = false;
  for (size_t n : bin.particles) {
    const abx_banked& b = bank[n];
    if (b.wgt > 0.)
      p1 = true;
    else if (b.wgt < 0.)
      n1 = true;
    if (b.wgt2 > 0.)
      p2 = true;
    else if (b.wgt2 < 0.)
      n2 = true;
    if (p1 && n1 && p2 && n2) break;
  }
  m1 = p1 && n1;
  m2 = p2 && n2;
}

// C(f, f_inv) and the bin sums of compute_averages (:343-364)
static void bin_sums(const std::vector<abx_banked>& bank, ExBin& bin) {
  auto C = [](double f, double f_inv) { return 1. / (2. * f * f_inv - 1.); };
  double sum_c = 0.;
  for (size_t i = 0; i < bin.particles.size(); i++) sum_c += C(bin.f[i], bin.f_inv[i]);
  double sum_c_wgt = 0., sum_c_wgt2 = 0.;
  for (size_t i = 0; i < bin.particles.size(); i++) {
    sum_c_wgt += C(bin.f[i], bin.f_inv[i]) * bank[bin.particles[i]].wgt;
    sum_c_wgt2 += C(bin.f[i], bin.f_inv[i]) * bank[bin.particles[i]].wgt2;
  }
  bin.sum_c = sum_c;
  bin.sum_c_wgt = sum_c_wgt;
  bin.sum_c_wgt2 = sum_c_wgt2;
}

// ExactMGCancelator::compute_averages (:294-365)
static void exact_averages(const Problem& P, const std::vector<abx_banked>& bank,
                           const std::vector<abx_parent_info>& par, const ExKey& key, int mat,
                           ExBin& bin) {
  bin.f.assign(bin.particles.size(), 0.);
  bin.f_inv.assign(bin.particles.size(), 0.);
  std::vector<std::pair<V3, size_t>> smp;
  uint64_t si = 0;
  for (uint32_t j = 0; j < (uint32_t)P.ex.n_samples; j++) {
    V3 r;
    if (!region_point(P, key, mat, &si, nullptr, r)) {
      bin.can_cancel = false;
      return;
    }
    size_t g = 0;
    if (P.ex.chi_matrix) {  // sample_point (:282-289): the 4th coordinate of i - 1
      const double xi = sobol_sample(si - 1, 3);
      const size_t gi = (size_t)std::floor(xi * (double)P.ex_groups[key.e].size());
      g = (size_t)P.ex_groups[key.e][gi];
    }
    smp.push_back({r, g});
  }
  const Material& m = P.materials[(size_t)mat];
  for (size_t i = 0; i < bin.particles.size(); i++) {
    const abx_parent_info& q = par[bin.particles[i]];
    const size_t g1 = settings_group(P, q.E1), g3 = settings_group(P, q.E3);
    double sf = 0., sfi = 0.;
    for (const auto& rs : smp) {
      const double f = exact_get_f(P, m, v3(q.r1), v3(q.u1), g1, g3, rs.first, rs.second, q.Esmp);
      if (f == 0.) {
        bin.can_cancel = false;
        return;
      }
      sf += f;
      sfi += 1. / f;
    }
    bin.f[i] = sf / (double)P.ex.n_samples;
    bin.f_inv[i] = sfi / (double)P.ex.n_samples;
  }
  bin_sums(bank, bin);
}

// BasicExactMGCancelator::get_averages(_sobol) (:243-407)
static void basic_averages(const Problem& P, const std::vector<abx_banked>& bank,
                           const std::vector<abx_parent_info>& par, const ExKey& key, int mat,
                           ExBin& bin, Pcg32& g) {
  bin.f.assign(bin.particles.size(), 0.);
  bin.f_inv.assign(bin.particles.size(), 0.);
  std::vector<V3> smp;
  uint64_t si = 0;
  for (uint32_t j = 0; j < (uint32_t)P.ex.n_samples; j++) {
    V3 r;
    if (!region_point(P, key, mat, P.ex.sobol ? &si : nullptr, &g, r)) {
      bin.can_cancel = false;
      return;
    }
    smp.push_back(r);
  }
  for (size_t i = 0; i < bin.particles.size(); i++) {
    const abx_parent_info& q = par[bin.particles[i]];
    double sf = 0., sfi = 0.;
    for (const V3& r : smp) {
      const double f = basic_get_f(r, v3(q.r1), q.Esmp);
      sf += f;
      sfi += 1. / f;
    }
    bin.f[i] = sf / (double)P.ex.n_samples;
    bin.f_inv[i] = sfi / (double)P.ex.n_samples;
  }
  if (P.ex.beta == ABX_BETA_AVERAGE_G) bin_sums(bank, bin);
}

// BasicExactMGCancelator::get_min_f (:218-241)
static double basic_min_f(const Problem& P, const ExKey& key, const V3& rp, double Esmp) {
  const double Xl = P.ex_low.x + (double)key.i * P.ex_pitch[0], Xh = Xl + P.ex_pitch[0];
  const double Yl = P.ex_low.y + (double)key.j * P.ex_pitch[1], Yh = Yl + P.ex_pitch[1];
  const double Zl = P.ex_low.z + (double)key.k * P.ex_pitch[2], Zh = Zl + P.ex_pitch[2];
  const double f1 = basic_get_f({Xh, Yh, Zh}, rp, Esmp), f2 = basic_get_f({Xh, Yh, Zl}, rp, Esmp);
  const double f3 = basic_get_f({Xh, Yl, Zh}, rp, Esmp), f4 = basic_get_f({Xh, Yl, Zl}, rp, Esmp);
  const double f5 = basic_get_f({Xl, Yh, Zh}, rp, Esmp), f6 = basic_get_f({Xl, Yh, Zl}, rp, Esmp);
  const double f7 = basic_get_f({Xl, Yl, Zh}, rp, Esmp), f8 = basic_get_f({Xl, Yl, Zl}, rp, Esmp);
  return std::min(std::min(std::min(f1, f2), std::min(f3, f4)),
                  std::min(std::min(f5, f6), std::min(f7, f8)));
}

// cancel_bin of either cancelator (exact :367-403, basic :438-468)
static void cancel_region(const Problem& P, std::vector<abx_banked>& bank,
                          const std::vector<abx_parent_info>& par, const ExKey& key, int mat,
                          ExBin& bin, bool first) {
  const bool basic = P.exact == ABX_BASIC_EXACT_MG;
  for (size_t i = 0; i < bin.particles.size(); i++) {
    abx_banked& b = bank[bin.particles[i]];
    const abx_parent_info& q = par[bin.particles[i]];
    const double wgt = first ? b.wgt : b.wgt2;
    if (basic && wgt == 0.) return;
    // get_beta
    double B = 0.;
    if (bin.can_cancel) {
      if (!basic || P.ex.beta == ABX_BETA_AVERAGE_G) {
        const double S = (first ? bin.sum_c_wgt : bin.sum_c_wgt2) / (1. + bin.sum_c);
        const double f = bin.f[i], f_inv = bin.f_inv[i];
        B = f * (1. / (2. * f * f_inv - 1.)) * (1. - (S / wgt));
      } else if (P.ex.beta == ABX_BETA_MINIMUM) {
        B = basic_min_f(P, key, v3(q.r1), q.Esmp);
      } else if (P.ex.beta == ABX_BETA_AVERAGE_F) {
        const double N = (double)bin.particles.size();
        const double W = first ? bin.W : bin.W2;
        B = bin.f[i] * (1. - ((W) / ((N + 1.) * wgt)));
      }
    }
    double f;
    if (basic) {
      f = basic_get_f(v3(b.r), v3(q.r1), q.Esmp);
    } else {
      f = exact_get_f(P, P.materials[(size_t)mat], v3(q.r1), v3(q.u1), settings_group(P, q.E1), settings_group(P, q.E3),
                      v3(b.r), settings_group(P, b.E), q.Esmp);
    }
    const double P_p = (f - B) / f;
    const double P_u = B / f;
    if (std::isinf(P_u) || std::isinf(P_p) || std::isnan(P_u) || std::isnan(P_p)) return;
    if (first) {
      bin.uniform_wgt += b.wgt * P_u;
      b.wgt *= P_p;
    } else {
      bin.uniform_wgt2 += b.wgt2 * P_u;
      b.wgt2 *= P_p;
    }
  }
}

// perform_cancellation + get_new_particles (exact :405-605, basic
// :471-611) with settings::rng; the new particles are appended to the bank.
static void exact_cancellation(Problem& P, std::vector<abx_banked>& bank,
                               const std::vector<abx_parent_info>& par) {
  if (P.exact == ABX_BASIC_EXACT_MG && P.ex.beta == ABX_BETA_ZERO) return;
  ExBins bins = exact_bins(P, bank, par);
  if (bins.empty()) return;
  const bool basic = P.exact == ABX_BASIC_EXACT_MG;
  // basic: each region's rng stream is settings::rng advanced past the
  // draws every earlier region may take (:475-489)
  uint64_t seed_advance = 0;
  if (basic && P.ex.beta != ABX_BETA_MINIMUM) {
    const uint64_t max_rn = (uint64_t)P.ex.n_samples * 100u *
                            (P.ex.beta == ABX_BETA_AVERAGE_G ? 2u : 1u);
    for (auto& kv : bins) {
      kv.second.second.rng_seed_advance = seed_advance;
      seed_advance += kv.second.second.particles.size() * max_rn;
    }
  }
  for (auto& kv : bins) {
    const ExKey& key = kv.second.first;
    ExBin& bin = kv.second.second;
    const int mat = std::get<4>(kv.first);
    Pcg32 local = P.grng;
    local.advance(bin.rng_seed_advance);
    if (bin.particles.size() > 1) {
      bool m1, m2;
      mixed_signs(bank, bin, m1, m2);
      if (m1 || m2) {
        if (!basic)
          exact_averages(P, bank, par, key, mat, bin);
        else if (P.ex.beta == ABX_BETA_AVERAGE_F || P.ex.beta == ABX_BETA_AVERAGE_G)
          basic_averages(P, bank, par, key, mat, bin, local);
      }
      if (m1) cancel_region(P, bank, par, key, mat, bin, true);
      if (m2) cancel_region(P, bank, par, key, mat, bin, false);
      bin.particles.clear();
      bin.f.clear();
      bin.f_inv.clear();
      bin.sum_c = bin.sum_c_wgt = bin.sum_c_wgt2 = 0.;
    }
  }
  if (basic && P.ex.beta != ABX_BETA_MINIMUM) P.grng.advance(seed_advance);
  // get_new_particles
  std::vector<abx_banked> uni;
  for (auto& kv : bins) {
    const ExKey& key = kv.second.first;
    ExBin& bin = kv.second.second;
    const int mat = std::get<4>(kv.first);
    const uint32_t N =
        (uint32_t)std::ceil(std::max(std::abs(bin.uniform_wgt), std::abs(bin.uniform_wgt2)));
    if (N > 0) {
      const double w = bin.uniform_wgt / N, w2 = bin.uniform_wgt2 / N;
      const Material& m = P.materials[(size_t)mat];
      for (uint32_t i = 0; i < N; i++) {
        V3 r;
        if (!region_point(P, key, mat, nullptr, &P.grng, r))
          fatal(ABX_E_INVALID, "Couldn't sample position for uniform particle.");
        abx_banked u{};
        u.r[0] = r.x;
        u.r[1] = r.y;
        u.r[2] = r.z;
        if (!basic) {
          size_t e_index = 0;
          if (P.ex.chi_matrix) {
            const double xi_E = RNG::rand(P.grng);
            const size_t gi = (size_t)std::floor(xi_E * (double)P.ex_groups[key.e].size());
            e_index = (size_t)P.ex_groups[key.e][gi];
          } else {
            e_index = (size_t)RNG::discrete(P.grng, m.chi[0].data(), m.chi[0].data() + m.chi[0].size());
          }
          u.E = 0.5 * (P.ebounds[e_index] + P.ebounds[e_index + 1]);
          // Direction u_smp(2 rand - 1, 2 pi rand) (:573): the order of the
          // two draws is the compiler's; GCC, the reference's usual
          // compiler, evaluates the arguments right to left - phi first
          const double phi = 2. * PI * RNG::rand(P.grng);
          const double mu = 2. * RNG::rand(P.grng) - 1.;
          const V3 d = direction_mu_phi(mu, phi);
          u.u[0] = d.x;
          u.u[1] = d.y;
          u.u[2] = d.z;
        } else {
          // MGNuclide::sample_fission(0, (0, 0, 1), 0, 0, rng) (src/mg_nuclide.cpp:504-544)
          const auto& chi = m.chi[0];
          const size_t ei = (size_t)RNG::discrete(P.grng, chi.data(), chi.data() + chi.size());
          u.E = 0.5 * (P.ebounds[ei] + P.ebounds[ei + 1]);
          const double mu = 2. * RNG::rand(P.grng) - 1.;
          const double phi = 2. * PI * RNG::rand(P.grng);
          const V3 d = rotate_direction({0., 0., 1.}, mu, phi);
          u.u[0] = d.x;
          u.u[1] = d.y;
          u.u[2] = d.z;
          (void)RNG::rand(P.grng);  // the delayed-neutron test, P_delayed = 0
        }
        u.wgt = w;
        u.wgt2 = w2;
        uni.push_back(u);
      }
    }
    bin.uniform_wgt = bin.uniform_wgt2 = 0.;
  }
  bank.insert(bank.end(), uni.begin(), uni.end());
}

// ApproximateMeshCancelator::add_particle + perform_cancellation
// (src/approximate_mesh_cancelator.cpp:97-190): bin by (i, j, k, e), and in
// every bin holding more than one particle with weights of both signs set
// each weight to the bin mean (w and w2 independently).  Bins keep the bank
// order, so the sums run in bank order.
static void approximate_cancellation(const Problem& P, std::vector<abx_banked>& bank) {
  std::map<int, std::vector<size_t>> bins;
  for (size_t n = 0; n < bank.size(); n++) {
    const abx_banked& p = bank[n];
    int i = (int)std::floor((p.r[0] - P.c_low.x) / P.c_dx);
    int j = (int)std::floor((p.r[1] - P.c_low.y) / P.c_dy);
    int k = (int)std::floor((p.r[2] - P.c_low.z) / P.c_dz);
    int l = -1;
    if (!P.c_edges.empty()) {
      for (size_t e = 0; e < P.c_edges.size() - 1; e++)
        if (P.c_edges[e] <= p.E && p.E <= P.c_edges[e + 1]) {
          l = (int)e;
          break;
        }
    } else {
      l = 0;
    }
    if (i < 0 || j < 0 || k < 0 || l < 0) continue;
    if (i >= (int)P.c_shape[0] || j >= (int)P.c_shape[1] || k >= (int)P.c_shape[2] ||
        l >= (int)P.c_shape[3])
      continue;
    const int key = l + (int)P.c_shape[3] * (k + (int)P.c_shape[2] * (j + (int)P.c_shape[1] * i));
    bins[key].push_back(n);
  }
  for (auto& kv : bins) {
    auto& bin = kv.second;
    if (bin.size() <= 1) continue;
    bool pos1 = false, neg1 = false, pos2 = false, neg2 = false;
    double sum1 = 0., sum2 = 0.;
    for (size_t n : bin) {
      const abx_banked& p = bank[n];
      if (p.wgt > 0.)
        pos1 = true;
      else if (p.wgt < 0.)
        neg1 = true;
      sum1 += p.wgt;
      if (p.wgt2 > 0.)
        pos2 = true;
      else if (p.wgt2 < 0.)
        neg2 = true;
      sum2 += p.wgt2;
    }
    const double N = (double)bin.size();
    const double avg1 = sum1 / N, avg2 = sum2 / N;
    for (size_t n : bin) {
      if (pos1 && neg1) bank[n].wgt = avg1;
      if (pos2 && neg2) bank[n].wgt2 = avg2;
    }
  }
}

// FixedSource::run, ModifiedFixedSource::run and PowerIterator::run
static void run(Problem& P, int ngen, int nignored, int nthreads,
                double* kcol_hist, abx_run_stats& S) {
  const uint64_t N = (uint64_t)P.st.nparticles;
  const double W = (double)N;
  uint64_t histories_counter = 0;
  std::vector<abx_particle> bank;
  double kcol = 1.0;
  bool converged = P.st.mode != ABX_K_EIGENVALUE || nignored == 0;
  S = abx_run_stats{};
  double seconds = 0;
  // a run starts its statistics afresh (as abx_run does), and
  // settings::rng from the seed (settings::initialize_global_rng,
  // src/settings.cpp:116-118)
  P.grng.seed(P.st.seed);
  P.gen = 0;
  P.k_col_avg = P.k_col_var = P.k_abs_avg = P.k_abs_var = P.k_trk_avg = P.k_trk_var = 0;
  P.k_tot_avg = P.k_tot_var = P.leak_avg = P.leak_var = P.mig_avg = P.mig_var = 0;
  for (auto& v : P.gen_vals) v.clear();
  for (auto& t : P.tallies) {
    t.g = 0;
    std::fill(t.avg.begin(), t.avg.end(), 0.);
    std::fill(t.var.begin(), t.var.end(), 0.);
  }
  if (P.st.mode == ABX_K_EIGENVALUE) {
    bank.resize(N);
    for (uint64_t i = 0; i < N; i++) bank[i] = sample_one(P, histories_counter++);
  }
  uint64_t global_counter = histories_counter;
  for (int g = 1; g <= ngen; g++) {
    tally_clear(P);
    Scores sc;
    abx_diag dg{};
    std::vector<abx_banked> next;
    if (P.st.mode == ABX_FIXED_SOURCE || P.st.mode == ABX_MODIFIED_FIXED_SOURCE) {
      bank.resize(N);
      for (uint64_t i = 0; i < N; i++) bank[i] = sample_one(P, histories_counter++);
      double t0 = omp_get_wtime();
      transport(P, bank.data(), bank.size(), kcol, true, nthreads, next, sc, dg,
                nullptr, 0, nullptr);
      while (P.st.mode == ABX_MODIFIED_FIXED_SOURCE && !next.empty()) {
        if (P.cancel) approximate_cancellation(P, next);
        bank.clear();
        for (auto& b : next) {  // src/modified_fixed_source.cpp:88-93
          abx_particle a{};
          for (int k = 0; k < 3; k++) {
            a.r[k] = b.r[k];
            a.u[k] = b.u[k];
          }
          a.E = b.E;
          a.wgt = b.wgt;
          a.wgt2 = b.wgt2;
          a.history_id = histories_counter++;
          a.family_id = 0;
          Pcg32 r;
          r.seed(P.st.seed);
          r.advance(P.st.stride * a.history_id);
          a.rng_state = r.state;
          bank.push_back(a);
        }
        next.clear();
        transport(P, bank.data(), bank.size(), kcol, true, nthreads, next, sc,
                  dg, nullptr, 0, nullptr);
      }
      seconds += omp_get_wtime() - t0;
      double kc = sc.k_col / W;
      if (kcol_hist) kcol_hist[g - 1] = kc;
      push_gen_values(P, sc, W);
      record_generation(P, sc, W);
    } else {
      double t0 = omp_get_wtime();
      transport(P, bank.data(), bank.size(), kcol, converged, nthreads, next, sc,
                dg, nullptr, 0, nullptr);
      seconds += omp_get_wtime() - t0;
      if (next.empty()) fatal(ABX_E_INVALID, "No fission neutrons were produced.");
      double kc = sc.k_col / W;
      if (kcol_hist) kcol_hist[g - 1] = kc;
      push_gen_values(P, sc, W);
      if (P.cancel) approximate_cancellation(P, next);
      if (P.exact) exact_cancellation(P, next, P.last_parents);
      // normalize_weights (src/power_iterator.cpp:538-586)
      double W_pos = 0., W_neg = 0.;
      for (auto& b : next) {
        if (b.wgt > 0.)
          W_pos += b.wgt;
        else
          W_neg -= b.wgt;
      }
      double Wn = W_pos - W_neg;
      double w_per_part = (double)P.st.nparticles / Wn;
      for (auto& b : next) b.wgt *= w_per_part;
      if (converged) record_generation(P, sc, W);
      // rebank (:394-404)
      histories_counter = global_counter;
      bank.clear();
      for (auto& b : next) {
        abx_particle a{};
        for (int k = 0; k < 3; k++) {
          a.r[k] = b.r[k];
          a.u[k] = b.u[k];
        }
        a.E = b.E;
        a.wgt = b.wgt;
        a.wgt2 = 0.;
        a.history_id = histories_counter++;
        a.family_id = b.family_id;
        Pcg32 r;
        r.seed(P.st.seed);
        r.advance(P.st.stride * a.history_id);
        a.rng_state = r.state;
        bank.push_back(a);
      }
      global_counter += next.size();
      kcol = kc;
      if (g == nignored) converged = true;
    }
    add_diag(S.diag, dg);
  }
  P.last_bank = bank;
  S.kcol_avg = P.k_col_avg;
  S.kcol_err = std::sqrt(P.k_col_var / (double)P.gen);
  S.kabs_avg = P.k_abs_avg;
  S.kabs_err = std::sqrt(P.k_abs_var / (double)P.gen);
  S.ktot_avg = P.k_tot_avg;
  S.ktot_err = std::sqrt(P.k_tot_var / (double)P.gen);
  S.ktrk_avg = P.k_trk_avg;
  S.ktrk_err = std::sqrt(P.k_trk_var / (double)P.gen);
  S.leak_avg = P.leak_avg;
  S.leak_err = std::sqrt(P.leak_var / (double)P.gen);
  S.mig_avg = P.mig_avg;
  S.mig_err = std::sqrt(P.mig_var / (double)P.gen);
  S.active_generations = P.gen;
  S.transport_seconds = seconds;
}

}  // namespace orc

using namespace orc;

struct orc_problem {
  Problem P;
};

static thread_local std::string g_last;

extern "C" {

const char* orc_last_error(void) { return g_last.c_str(); }

double orc_log(int portable, double x) { return portable ? dm_log(x) : std::log(x); }
double orc_exp(int portable, double x) { return portable ? dm_exp(x) : std::exp(x); }
double orc_sin(int portable, double x) { return portable ? dm_sin(x) : std::sin(x); }
double orc_cos(int portable, double x) { return portable ? dm_cos(x) : std::cos(x); }

#define ORC_TRY(...)                      \
  try {                                   \
    g_portable = p->P.portable;           \
    __VA_ARGS__;                          \
    return ABX_OK;                        \
  } catch (const Fatal& f) {              \
    g_last = f.msg;                       \
    return f.code;                        \
  } catch (const std::exception& e) {     \
    g_last = e.what();                    \
    return ABX_E_INVALID;                 \
  }

int orc_create(const abx_problem_desc* desc, int portable_math,
               orc_problem** out) {
  try {
    auto* p = new orc_problem;
    p->P.portable = portable_math != 0;
    g_portable = p->P.portable;
    build(p->P, desc);
    p->P.grng.seed(p->P.st.seed);
    *out = p;
    return ABX_OK;
  } catch (const Fatal& f) {
    g_last = f.msg;
    return f.code;
  }
}

void orc_destroy(orc_problem* p) { delete p; }

int orc_sample_sources(orc_problem* p, uint64_t first, size_t n,
                       abx_particle* out) {
  ORC_TRY(for (size_t i = 0; i < n; i++) out[i] = sample_one(p->P, first + i));
}

int orc_transport(orc_problem* p, const abx_particle* bank, size_t n,
                  double kcol, int converged, int nthreads,
                  abx_banked* fission_out, size_t cap, size_t* n_out,
                  abx_scores* scores, abx_diag* diag, abx_trace_rec* trace,
                  int cap_per_history, int32_t* n_rec) {
  ORC_TRY({
    std::vector<abx_banked> fo;
    Scores sc;
    abx_diag dg{};
    transport(p->P, bank, n, kcol, converged != 0, nthreads, fo, sc, dg, trace,
              cap_per_history, n_rec);
    if (n_out) *n_out = fo.size();
    if (fo.size() > cap && fission_out) fatal(ABX_E_CAPACITY, "fission bank capacity");
    if (fission_out) std::copy(fo.begin(), fo.end(), fission_out);
    p->P.gen_scores.k_col += sc.k_col;
    if (scores) {
      scores->k_col = sc.k_col;
      scores->k_abs = sc.k_abs;
      scores->k_trk = sc.k_trk;
      scores->k_tot = sc.k_tot;
      scores->leakage = sc.leak;
      scores->mig = sc.mig;
    }
    if (diag) *diag = dg;
  });
}

int orc_tally_size(orc_problem* p, int tally, size_t* n_bins) {
  ORC_TRY({
    if (tally < 0 || tally >= (int)p->P.tallies.size()) fatal(ABX_E_INVALID, "tally index");
    *n_bins = p->P.tallies[(size_t)tally].gen.size();
  });
}

int orc_tally_fetch(orc_problem* p, int tally, double* gen_out) {
  ORC_TRY({
    if (tally < 0 || tally >= (int)p->P.tallies.size()) fatal(ABX_E_INVALID, "tally index");
    auto& t = p->P.tallies[(size_t)tally];
    std::copy(t.gen.begin(), t.gen.end(), gen_out);
  });
}

int orc_tally_clear(orc_problem* p) { ORC_TRY(tally_clear(p->P)); }

// MeshTally::record_generation(multiplier) of every mesh tally
// (src/mesh_tally.cpp:121-150), for a driver that calls transport itself.
int orc_tally_record(orc_problem* p, double multiplier) {
  ORC_TRY({
    for (auto& t : p->P.tallies) {
      t.g++;
      const double dg = (double)t.g;
      for (size_t i = 0; i < t.gen.size(); i++) {
        double old_avg = t.avg[i];
        double val = t.gen[i] * multiplier;
        double avg = old_avg + (val - old_avg) / dg;
        t.avg[i] = avg;
        double var = t.var[i];
        var = var + (((val - old_avg) * (val - avg) - (var)) / dg);
        t.var[i] = var;
      }
    }
  });
}

int orc_parents_fetch(orc_problem* p, abx_parent_info* out, size_t cap, size_t* n) {
  ORC_TRY({
    *n = p->P.last_parents.size();
    if (*n > cap) fatal(ABX_E_CAPACITY, "parent buffer too small");
    std::copy(p->P.last_parents.begin(), p->P.last_parents.end(), out);
  });
}

int orc_exact_cancel(orc_problem* p, const abx_banked* bank, const abx_parent_info* par, size_t n,
                     abx_banked* out, size_t cap, size_t* n_out) {
  ORC_TRY({
    if (!p->P.exact) fatal(ABX_E_INVALID, "no exact cancelator in this problem");
    std::vector<abx_banked> b(bank, bank + n);
    std::vector<abx_parent_info> q(par, par + n);
    exact_cancellation(p->P, b, q);
    *n_out = b.size();
    if (b.size() > cap) fatal(ABX_E_CAPACITY, "cancelled bank buffer too small");
    std::copy(b.begin(), b.end(), out);
  });
}

int orc_rng_reset(orc_problem* p) { ORC_TRY(p->P.grng.seed(p->P.st.seed)); }

uint64_t orc_sobol_matrix(unsigned dim, unsigned k) { return sobol_matrices()[dim * 52 + k]; }

double orc_sobol(uint64_t index, unsigned dim) { return sobol_sample(index, dim); }

int orc_run(orc_problem* p, int ngenerations, int nignored, int nthreads,
            double* kcol_hist, abx_run_stats* stats) {
  ORC_TRY(run(p->P, ngenerations, nignored, nthreads, kcol_hist, *stats));
}

int orc_tally_stats(orc_problem* p, int tally, double* avg, double* std_err) {
  ORC_TRY({
    if (tally < 0 || tally >= (int)p->P.tallies.size()) fatal(ABX_E_INVALID, "tally index");
    auto& t = p->P.tallies[(size_t)tally];
    for (size_t i = 0; i < t.avg.size(); i++) {
      avg[i] = t.avg[i];
      std_err[i] = std::sqrt(t.var[i] / (double)t.g);
    }
  });
}

int orc_gen_values(orc_problem* p, int quantity, double* out, size_t cap, size_t* n) {
  ORC_TRY({
    if (quantity < 0 || quantity > 5) fatal(ABX_E_INVALID, "quantity");
    const auto& v = p->P.gen_vals[quantity];
    *n = v.size();
    if (v.size() > cap) fatal(ABX_E_CAPACITY, "per-generation buffer too small");
    std::copy(v.begin(), v.end(), out);
  });
}

int orc_bank_size(orc_problem* p, size_t* n) { ORC_TRY(*n = p->P.last_bank.size()); }

int orc_bank_fetch(orc_problem* p, double* out9) {
  ORC_TRY({
    const auto& b = p->P.last_bank;
    for (size_t i = 0; i < b.size(); i++) {
      double* o = out9 + 9 * i;
      o[0] = b[i].r[0];
      o[1] = b[i].r[1];
      o[2] = b[i].r[2];
      o[3] = b[i].u[0];
      o[4] = b[i].u[1];
      o[5] = b[i].u[2];
      o[6] = b[i].E;
      o[7] = b[i].wgt;
      o[8] = b[i].wgt2;
    }
  });
}

int orc_locate(orc_problem* p, const double* r, const double* u, size_t n,
               int32_t* cell, int32_t* material) {
  ORC_TRY({
    for (size_t i = 0; i < n; i++) {
      Tracker t(&p->P, {r[3 * i], r[3 * i + 1], r[3 * i + 2]},
                {u[3 * i], u[3 * i + 1], u[3 * i + 2]});
      cell[i] = t.cell;
      material[i] = t.material();
    }
  });
}

int orc_nearest_boundary(orc_problem* p, const double* r, const double* u,
                         size_t n, double* dist, int32_t* surface,
                         int32_t* btype) {
  ORC_TRY({
    for (size_t i = 0; i < n; i++) {
      Tracker t(&p->P, {r[3 * i], r[3 * i + 1], r[3 * i + 2]},
                {u[3 * i], u[3 * i + 1], u[3 * i + 2]});
      if (t.is_lost()) {
        dist[i] = -1;
        surface[i] = -1;
        btype[i] = -1;
        continue;
      }
      BoundaryInfo b = t.get_nearest_boundary();
      dist[i] = b.distance;
      surface[i] = b.surface_index;
      btype[i] = b.btype;
    }
  });
}

int orc_cancel(orc_problem* p, abx_banked* bank, size_t n) {
  ORC_TRY({
    if (!p->P.cancel) fatal(ABX_E_INVALID, "no cancellation mesh");
    std::vector<abx_banked> b(bank, bank + n);
    approximate_cancellation(p->P, b);
    std::copy(b.begin(), b.end(), bank);
  });
}

int orc_majorant(orc_problem* p, double* maj) {
  ORC_TRY(std::copy(p->P.majorant.begin(), p->P.majorant.end(), maj));
}

}  // extern "C"

// ---- RNG probes (golden-vector tests) ---------------------------------------
extern "C" {
void orc_rng_stream(uint64_t seed, uint64_t stride, uint64_t h, int n, uint32_t* out) {
  Pcg32 g;
  g.seed(seed);
  g.advance(stride * h);
  for (int i = 0; i < n; i++) out[i] = g();
}
void orc_rng_rand(uint64_t seed, uint64_t stride, uint64_t h, int n, double* out) {
  Pcg32 g;
  g.seed(seed);
  g.advance(stride * h);
  for (int i = 0; i < n; i++) out[i] = RNG::rand(g);
}
void orc_rng_exponential(int portable, uint64_t seed, uint64_t stride, uint64_t h, int n,
                         double lambda, double* out) {
  g_portable = portable != 0;
  Pcg32 g;
  g.seed(seed);
  g.advance(stride * h);
  for (int i = 0; i < n; i++) out[i] = RNG::exponential(g, lambda);
}
uint32_t orc_rng_discrete(uint64_t seed, uint64_t stride, uint64_t h, const double* w, int nw,
                          int n, int32_t* out) {
  Pcg32 g;
  g.seed(seed);
  g.advance(stride * h);
  for (int i = 0; i < n; i++) out[i] = RNG::discrete(g, w, w + nw);
  return g();
}
}
