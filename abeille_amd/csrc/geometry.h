// geometry.h - CSG navigation on the flattened tables (device code).
//
// Restates, on SoA tables and a fixed-size per-particle cursor:
//  * Surface::{sign,distance,norm}: src/xplane.cpp:32-56 (y/z alike),
//    src/plane.cpp:33-60, src/{x,y,z}cylinder.cpp (e.g. src/zcylinder.cpp:33-80),
//    src/sphere.cpp:33-80 - same operand order, same 1e-12 tolerances;
//  * Cell::is_inside / distance_to_boundary(_condition): src/cell.cpp:71-193;
//  * RectLattice::get_tile / distance_to_tile_boundary / tile_center:
//    src/rect_lattice.cpp:209-310; HexLattice's: src/hex_lattice.cpp:245-452;
//  * CellUniverse / RectLattice / HexLattice ::get_cell with the lily-pad
//    stack (src/hex_lattice.cpp:143-217) and
//    Tracker::{move,get_current,get_boundary_condition,get_nearest_boundary}:
//    include/simulation/tracker.hpp:41-372, src/cell_universe.cpp:72-150,
//    src/lattice.cpp:77-91.
//
// The reference's vector-of-pads becomes a bounded cursor: pads (type,
// index, frame, tile) and one coordinate frame per nesting level.  Pads of
// one level share a frame because the reference moves every pad by the same
// d*u (tracker.hpp:76-85), so their r_local stay bit-identical; local
// positions are advanced incrementally exactly like the reference.
#pragma once

#include <math.h>

#include "../../include/abx.h"
#include "devtypes.h"

struct V3 {
  double x, y, z;
};

ABX_HD V3 vadd(const V3& a, const V3& b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
ABX_HD V3 vsub(const V3& a, const V3& b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
ABX_HD V3 vscale(const V3& v, double d) { return V3{v.x * d, v.y * d, v.z * d}; }
ABX_HD double vdot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ABX_HD double vnorm(const V3& a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Direction(x, y, z): normalising constructor (include/utils/direction.hpp:37-42)
ABX_HD V3 vdir(double x, double y, double z) {
  const double m = sqrt(x * x + y * y + z * z);
  return V3{x / m, y / m, z / m};
}
ABX_HD double dabs(double x) { return fabs(x); }

// ---------------------------------------------------------------- surfaces
// EXT builds also evaluate Cylinder (src/cylinder.cpp:59-117) off the
// coordinate axes, type ABX_CYL: the quadric alpha x^2 + beta y^2 +
// gamma z^2 = R^2 about (x0, y0, z0), its p = {x0, y0, z0, R} and alpha,
// beta, gamma in p[0..2] of a companion entry pad_[0] entries further on
// in the same table (appended after the deck's surfaces).
ABX_HD const DSurf& surf_aux(const DSurf& s) { return (&s)[s.pad_[0]]; }

template <bool EXT>
ABX_HD V3 surf_norm(const DSurf& s, const V3& r) {
  if (EXT && s.type == ABX_CYL) {
    const DSurf& q = surf_aux(s);
    return vdir(q.p[0] * (r.x - s.p[0]), q.p[1] * (r.y - s.p[1]), q.p[2] * (r.z - s.p[2]));
  }
  switch (s.type) {
    case 0: return vdir(1., 0., 0.);
    case 1: return vdir(0., 1., 0.);
    case 2: return vdir(0., 0., 1.);
    case 3: return vdir(s.p[0], s.p[1], s.p[2]);
    case 4: return vdir(0., r.y - s.p[0], r.z - s.p[1]);
    case 5: return vdir(r.x - s.p[0], 0., r.z - s.p[1]);
    case 6: return vdir(r.x - s.p[0], r.y - s.p[1], 0.);
    default: return vdir(r.x - s.p[0], r.y - s.p[1], r.z - s.p[2]);
  }
}

// A point on the surface (|eval| <= 1e-12) takes the sign of u . n.
template <bool EXT>
ABX_COLD int surf_sign_on(const DSurf& s, V3 r, V3 u) {
  return vdot(u, surf_norm<EXT>(s, r)) > 0. ? 1 : -1;
}

template <bool EXT>
ABX_HD int surf_sign(const DSurf& s, const V3& r, const V3& u) {
  double eval;
  if (EXT && s.type == ABX_CYL) {
    const DSurf& q = surf_aux(s);
    const double x = r.x - s.p[0], y = r.y - s.p[1], z = r.z - s.p[2];
    eval = q.p[0] * x * x + q.p[1] * y * y + q.p[2] * z * z - s.p[3] * s.p[3];
  } else {
  switch (s.type) {
    case 0: eval = r.x - s.p[0]; break;
    case 1: eval = r.y - s.p[0]; break;
    case 2: eval = r.z - s.p[0]; break;
    case 3: eval = s.p[0] * r.x + s.p[1] * r.y + s.p[2] * r.z - s.p[3]; break;
    case 4: {
      const double y = r.y - s.p[0], z = r.z - s.p[1];
      eval = y * y + z * z - s.p[2] * s.p[2];
      break;
    }
    case 5: {
      const double x = r.x - s.p[0], z = r.z - s.p[1];
      eval = x * x + z * z - s.p[2] * s.p[2];
      break;
    }
    case 6: {
      const double x = r.x - s.p[0], y = r.y - s.p[1];
      eval = y * y + x * x - s.p[2] * s.p[2];
      break;
    }
    default: {
      const double x = r.x - s.p[0], y = r.y - s.p[1], z = r.z - s.p[2];
      eval = (x * x) + (y * y) + (z * z) - s.p[3] * s.p[3];
      break;
    }
  }
  }
  if (eval > ABX_SURFACE_COINCIDENT) return 1;
  if (eval < -ABX_SURFACE_COINCIDENT) return -1;
  return surf_sign_on<EXT>(s, r, u);
}

ABX_HD double plane_axis_distance(double x0, double r0, double u0, bool on) {
  const double diff = x0 - r0;
  if (on || dabs(diff) < ABX_SURFACE_COINCIDENT || u0 == 0.) return ABX_INF;
  if (diff / u0 < 0.) return ABX_INF;
  return diff / u0;
}

ABX_HD double quad_distance(double a, double k, double c, bool on) {
  const double quad = k * k - a * c;
  if (quad < 0.) return ABX_INF;
  if (on || dabs(c) < ABX_SURFACE_COINCIDENT) {
    if (k >= 0.) return ABX_INF;
    return (-k + sqrt(quad)) / a;
  }
  if (c < 0.) return (-k + sqrt(quad)) / a;
  const double d = (-k - sqrt(quad)) / a;
  return d < 0. ? ABX_INF : d;
}

template <bool EXT>
ABX_NOINL double surf_distance(const DSurf& s, const V3& r, const V3& u, bool on) {
  if (EXT && s.type == ABX_CYL) {
    const DSurf& q = surf_aux(s);
    const double a = q.p[0] * u.x * u.x + q.p[1] * u.y * u.y + q.p[2] * u.z * u.z;
    if (a == 0.) return ABX_INF;
    const double x = r.x - s.p[0], y = r.y - s.p[1], z = r.z - s.p[2];
    const double k = q.p[0] * x * u.x + q.p[1] * y * u.y + q.p[2] * z * u.z;
    const double c = q.p[0] * x * x + q.p[1] * y * y + q.p[2] * z * z - s.p[3] * s.p[3];
    return quad_distance(a, k, c, on);
  }
  switch (s.type) {
    case 0: return plane_axis_distance(s.p[0], r.x, u.x, on);
    case 1: return plane_axis_distance(s.p[0], r.y, u.y, on);
    case 2: return plane_axis_distance(s.p[0], r.z, u.z, on);
    case 3: {
      const double num = s.p[3] - s.p[0] * r.x - s.p[1] * r.y - s.p[2] * r.z;
      const double denom = s.p[0] * u.x + s.p[1] * u.y + s.p[2] * u.z;
      const double d = num / denom;
      if (on || dabs(d) < ABX_SURFACE_COINCIDENT || denom == 0.) return ABX_INF;
      return d < 0. ? ABX_INF : d;
    }
    case 4: {
      const double a = u.y * u.y + u.z * u.z;
      if (a == 0.) return ABX_INF;
      const double y = r.y - s.p[0], z = r.z - s.p[1];
      return quad_distance(a, y * u.y + z * u.z, y * y + z * z - s.p[2] * s.p[2], on);
    }
    case 5: {
      const double a = u.x * u.x + u.z * u.z;
      if (a == 0.) return ABX_INF;
      const double x = r.x - s.p[0], z = r.z - s.p[1];
      return quad_distance(a, x * u.x + z * u.z, x * x + z * z - s.p[2] * s.p[2], on);
    }
    case 6: {
      const double a = u.y * u.y + u.x * u.x;
      if (a == 0.) return ABX_INF;
      const double x = r.x - s.p[0], y = r.y - s.p[1];
      return quad_distance(a, y * u.y + x * u.x, y * y + x * x - s.p[2] * s.p[2], on);
    }
    default: {
      const double x = r.x - s.p[0], y = r.y - s.p[1], z = r.z - s.p[2];
      const double k = x * u.x + y * u.y + z * u.z;
      const double c = x * x + y * y + z * z - s.p[3] * s.p[3];
      const double quad = k * k - c;
      if (quad < 0.) return ABX_INF;
      if (on || dabs(c) < ABX_SURFACE_COINCIDENT) {
        if (k >= 0.) return ABX_INF;
        return -k + sqrt(quad);
      }
      if (c < 0.) return -k + sqrt(quad);
      const double d = -k - sqrt(quad);
      return d < 0. ? ABX_INF : d;
    }
  }
}

// ------------------------------------------------------------------- cells
ABX_HD int iabs(int x) { return x < 0 ? -x : x; }

template <bool EXT>
ABX_HD bool cell_inside(const DevProblem& P, int ci, const V3& r, const V3& u,
                        int32_t on_surf) {
  const DCell& c = P.cell[ci];
  const int32_t* rpn = P.rpn + c.rpn_off;
  if (c.simple) {
    for (int k = 0; k < c.rpn_len; k++) {
      const int32_t token = rpn[k];
      if (token == on_surf) continue;
      if (-token == on_surf) return false;
      const int sign = surf_sign<EXT>(P.surf[iabs(token) - 1], r, u);
      if ((sign > 0 && token < 0) || (sign < 0 && token > 0)) return false;
    }
    return true;
  }
  // general region: boolean stack as a bit-vector (depth <= 64)
  uint64_t st = 0;
  int i = -1;
  for (int k = 0; k < c.rpn_len; k++) {
    const int32_t token = rpn[k];
    if (token == 2147483647 - 4) {  // union
      const uint64_t a = (st >> (i - 1)) & 1u, b = (st >> i) & 1u;
      st = (st & ~(1ull << (i - 1))) | ((a | b) << (i - 1));
      i--;
    } else if (token == 2147483647 - 3) {  // intersection
      const uint64_t a = (st >> (i - 1)) & 1u, b = (st >> i) & 1u;
      st = (st & ~(1ull << (i - 1))) | ((a & b) << (i - 1));
      i--;
    } else if (token == 2147483647 - 2) {  // complement
      st ^= (1ull << i);
    } else {
      i++;
      bool v;
      if (token == on_surf)
        v = true;
      else if (-token == on_surf)
        v = false;
      else {
        const int sign = surf_sign<EXT>(P.surf[iabs(token) - 1], r, u);
        v = (sign > 0 && token > 0) || (sign < 0 && token < 0);
      }
      st = (st & ~(1ull << i)) | ((uint64_t)(v ? 1u : 0u) << i);
    }
  }
  if (i == 0) return (st & 1u) != 0;
  return true;
}

struct DistTok {
  double d;
  int32_t tok;
};

// Cell::distance_to_boundary / distance_to_boundary_condition
template <bool EXT>
ABX_NOINL DistTok cell_distance(const DevProblem& P, int ci, const V3& r, const V3& u,
                                int32_t on_surf, bool bc_only) {
  const DCell& c = P.cell[ci];
  DistTok res{ABX_INF, 0};
  if (bc_only && !c.has_bc) return res;
  const int32_t* rpn = P.rpn + c.rpn_off;
  for (int k = 0; k < c.rpn_len; k++) {
    const int32_t token = rpn[k];
    if (token >= 2147483647 - 4) continue;
    const bool coincident = iabs(token) == iabs(on_surf);
    const DSurf& s = P.surf[iabs(token) - 1];
    if (bc_only && s.bc == 2) continue;
    const double d = surf_distance<EXT>(s, r, u, coincident);
    if (d < res.d) {
      if (dabs(d - res.d) / res.d >= 1e-14) {
        res.d = d;
        res.tok = -token;
      }
    }
  }
  return res;
}

// ---------------------------------------------------------------- lattices
ABX_HD V3 tile_center(const DUniv& L, int nx, int ny, int nz) {
  return V3{((double)nx + 0.5) * L.Px + L.Xl, ((double)ny + 0.5) * L.Py + L.Yl,
            ((double)nz + 0.5) * L.Pz + L.Zl};
}

ABX_HD void get_tile(const DUniv& L, const V3& r, const V3& u, int t[3]) {
  int nx = (int)floor((r.x - L.Xl) * L.Pxi);
  int ny = (int)floor((r.y - L.Yl) * L.Pyi);
  int nz = (int)floor((r.z - L.Zl) * L.Pzi);
  const V3 c = tile_center(L, nx, ny, nz);
  const double xl = c.x - L.Px * 0.5;
  if (dabs(xl - r.x) < ABX_SURFACE_COINCIDENT && u.x < 0.) nx--;
  const double xh = c.x + L.Px * 0.5;
  if (dabs(xh - r.x) < ABX_SURFACE_COINCIDENT && u.x >= 0.) nx++;
  const double yl = c.y - L.Py * 0.5;
  if (dabs(yl - r.y) < ABX_SURFACE_COINCIDENT && u.y < 0.) ny--;
  const double yh = c.y + L.Py * 0.5;
  if (dabs(yh - r.y) < ABX_SURFACE_COINCIDENT && u.y >= 0.) ny++;
  const double zl = c.z - L.Pz * 0.5;
  if (dabs(zl - r.z) < ABX_SURFACE_COINCIDENT && u.z < 0.) nz--;
  const double zh = c.z + L.Pz * 0.5;
  if (dabs(zh - r.z) < ABX_SURFACE_COINCIDENT && u.z >= 0.) nz++;
  t[0] = nx;
  t[1] = ny;
  t[2] = nz;
}

ABX_HD double tile_distance(const DUniv& L, const V3& rl, const V3& u, const int32_t t[3]) {
  const V3 c = tile_center(L, t[0], t[1], t[2]);
  const V3 rt = vsub(rl, c);
  double dist = ABX_INF;
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
  const double tol = 100 * ABX_SURFACE_COINCIDENT;
  if (d_xl > 0. && d_xl < dist && dabs(diff_xl) > tol) dist = d_xl;
  if (d_xh > 0. && d_xh < dist && dabs(diff_xh) > tol) dist = d_xh;
  if (d_yl > 0. && d_yl < dist && dabs(diff_yl) > tol) dist = d_yl;
  if (d_yh > 0. && d_yh < dist && dabs(diff_yh) > tol) dist = d_yh;
  if (d_zl > 0. && d_zl < dist && dabs(diff_zl) > tol) dist = d_zl;
  if (d_zh > 0. && d_zh < dist && dabs(diff_zh) > tol) dist = d_zh;
  return dist;
}

// ------------------------------------------------------- hexagonal lattices
// HexLattice (src/hex_lattice.cpp), expression for expression (the 0.* and
// 1.* factors included) so that tiles and distances carry the host's bits.

// HexLattice::get_nearest_hex + get_tile (src/hex_lattice.cpp:260-308): the
// axial (q, r) of the nearest centre by cube rounding, the axial layer from
// z_low; the direction plays no part.  With `elems`, also the tile's element
// or -1 outside the rings / layers (get_ring, linear_index:
// src/hex_lattice.cpp:245-258, 333-345).
struct HexTile {
  int32_t q, r, z, elem;
};
ABX_HD HexTile hex_tile(const DUniv& L, const int32_t* elems, double px, double py, double pz) {
  const double pt = L.Px;
  double q, r, det;
  if (L.top == ABX_HEX_FLAT) {
    det = -pt * pt * L.s3;
    q = (pt / det) * (0. * px - 1. * py);
    r = (pt / det) * (-L.s3 * px + L.c3 * py);
  } else {
    det = -pt * pt * L.c6;
    q = (pt / det) * (L.s6 * px - L.c6 * py);
    r = (pt / det) * (-1. * px - 0. * py);
  }
  const double cx = q, cz = r, cy = -cx - cz;
  double rx = round(cx), ry = round(cy), rz = round(cz);
  const double dx = fabs(rx - cx), dy = fabs(ry - cy), dz = fabs(rz - cz);
  if (dx > dy && dx > dz)
    rx = -ry - rz;
  else if (dy > dx && dy > dz)
    ry = -rx - rz;
  else
    rz = -rx - ry;
  HexTile t;
  t.q = (int32_t)rx;
  t.r = (int32_t)rz;
  t.z = (int32_t)floor((pz - L.z_low) / L.Pz);
  t.elem = -1;
  if (elems) {
    const int32_t y = -t.q - t.r;
    const uint32_t aq = (uint32_t)abs(t.q), ay = (uint32_t)abs(y), ar = (uint32_t)abs(t.r);
    const uint32_t ring = (aq > ay ? aq : ay) > ar ? (aq > ay ? aq : ay) : ar;
    if (ring < (uint32_t)L.rings && t.z >= 0 && t.z < L.Nz) {
      const int mid = L.rings - 1;
      t.elem = elems[L.elems_off + (t.z * L.Ny + (t.r + mid)) * L.Nx + (t.q + mid)];
    }
  }
  return t;
}

// HexLattice::tile_center / get_hex_center (src/hex_lattice.cpp:310-331)
ABX_HD V3 hex_tile_center(const DUniv& L, int q, int r, int nz) {
  const double dq = (double)q, dr = (double)r;
  V3 c;
  if (L.top == ABX_HEX_FLAT) {
    c.x = L.Px * (L.c3 * dq + 1. * dr);
    c.y = L.Px * (L.s3 * dq + 0. * dr);
  } else {
    c.x = L.Px * (0. * dq + L.c6 * dr);
    c.y = L.Px * (1. * dq + L.s6 * dr);
  }
  c.z = ((double)nz + 0.5) * L.Pz + L.z_low;
  return c;
}

// HexLattice::distance_to_line (src/hex_lattice.cpp:435-452)
ABX_HD double hex_line(const V3& r, const V3& u, double x1, double y1, double x2, double y2) {
  const double A = y2 - y1;
  const double B = x1 - x2;
  const double D = (x2 - x1) * y1 - (y2 - y1) * x1;
  const double num = D - A * r.x - B * r.y;
  const double denom = A * u.x + B * u.y;
  const double d = num / denom;
  if (d < 0.) return HUGE_VAL;
  if (fabs(num) <= 100 * ABX_SURFACE_COINCIDENT * sqrt(A * A + B * B)) return HUGE_VAL;
  return d;
}

// HexLattice::distance_to_tile_boundary(_pointy / _flat)
// (src/hex_lattice.cpp:347-433): the reference's six lines (which are not
// the edges of the hexagon get_tile assigns) and two axial planes, in the
// frame of the tile centre taken from r_local as given.  Lines and planes
// within 100 * SURFACE_COINCIDENT of the point are skipped, RectLattice's
// tolerance (src/rect_lattice.cpp:268-281): the reference takes any
// positive distance, and since its lines are not tile edges a particle
// stopped on one can find it again ~1e-16 cm ahead and cross it for ever
// (src/surface_tracker.cpp:72-150).
ABX_HD double hex_tile_distance(const DUniv& L, double rx, double ry, double rz, double ux,
                                  double uy, double uz, int q, int r, int nz) {
  const V3 u{ux, uy, uz};
  const V3 rt = vsub(V3{rx, ry, rz}, hex_tile_center(L, q, r, nz));
  const double pt = L.Px;
  const bool pointy = L.top == ABX_HEX_POINTY;
  // the six lines as (x1, y1) -> (x2, y2) in the reference's order, from the
  // orientation's two points (a, b), (c, e)
  const double a = pointy ? 0. : pt / (2. * L.s6);
  const double b = pointy ? pt / (2. * L.c6) : pt / 2.;
  const double c = pointy ? pt / 2. : pt / (2. * L.c6);
  const double e = pointy ? pt / (2. * L.s6) : 0.;
  double d = ABX_INF;
  for (int k = 0; k < 6; k++) {
    double x1, y1, x2, y2;
    if (pointy) {
      switch (k) {
        case 0: x1 = a; y1 = b; x2 = c; y2 = b; break;
        case 1: x1 = c; y1 = e; x2 = c; y2 = -e; break;
        case 2: x1 = c; y1 = -e; x2 = a; y2 = -b; break;
        case 3: x1 = a; y1 = -b; x2 = -c; y2 = -e; break;
        case 4: x1 = -c; y1 = -e; x2 = -c; y2 = e; break;
        default: x1 = -c; y1 = e; x2 = a; y2 = b; break;
      }
    } else {
      switch (k) {
        case 0: x1 = a; y1 = b; x2 = c; y2 = b; break;
        case 1: x1 = c; y1 = e; x2 = a; y2 = -b; break;
        case 2: x1 = a; y1 = -b; x2 = -a; y2 = -b; break;
        case 3: x1 = -a; y1 = -b; x2 = -c; y2 = -e; break;
        case 4: x1 = -c; y1 = -e; x2 = -a; y2 = b; break;
        default: x1 = -a; y1 = b; x2 = a; y2 = b; break;
      }
    }
    const double dk = hex_line(rt, u, x1, y1, x2, y2);
    if (dk > 0. && dk < d) d = dk;
  }
  const double tol = 100 * ABX_SURFACE_COINCIDENT;
  const double dzl = (-L.Pz * 0.5 - rt.z) / u.z;
  const double dzu = (L.Pz * 0.5 - rt.z) / u.z;
  if (dzl > 0. && dzl < d && fabs(-L.Pz * 0.5 - rt.z) > tol) d = dzl;
  if (dzu > 0. && dzu < d && fabs(L.Pz * 0.5 - rt.z) > tol) d = dzu;
  return d;
}

// Lattice::get_tile / distance_to_tile_boundary of either kind, from a
// lattice pad's r_local (the frame the lattice was entered in).
template <bool EXT>
ABX_HD void lat_pad_tile(const DUniv& L, const V3& r, const V3& u, int t[3]) {
  if (EXT && L.type == ABX_HEX_LATTICE) {
    const HexTile h = hex_tile(L, nullptr, r.x, r.y, r.z);
    t[0] = h.q;
    t[1] = h.r;
    t[2] = h.z;
  } else {
    get_tile(L, r, u, t);
  }
}
template <bool EXT>
ABX_HD double lat_pad_distance(const DUniv& L, const V3& rl, const V3& u, const int32_t t[3]) {
  if (EXT && L.type == ABX_HEX_LATTICE)
    return hex_tile_distance(L, rl.x, rl.y, rl.z, u.x, u.y, u.z, t[0], t[1], t[2]);
  return tile_distance(L, rl, u, t);
}

// ------------------------------------------------------------------ cursor
enum { PAD_U = 0, PAD_L = 1, PAD_C = 2 };

// pad word: type (2 bits) << 30 | level (3 bits) << 27 | index (27 bits)
ABX_HD int32_t pad_word(int type, int lvl, int idx) {
  return (int32_t)(((uint32_t)type << 30) | ((uint32_t)lvl << 27) | (uint32_t)idx);
}
ABX_HD int pad_type(int32_t w) { return (int)(((uint32_t)w) >> 30); }
ABX_HD int pad_lvl(int32_t w) { return (int)((((uint32_t)w) >> 27) & 7u); }
ABX_HD int pad_idx(int32_t w) { return (int)(((uint32_t)w) & 0x07ffffffu); }
ABX_HD bool tile_eq(const int32_t a[3], const int b[3]) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// The cursor proper.  Pads and frames live outside the struct (per-thread
// slices of LDS on the GPU, a small local buffer on the host) at a stride,
// so that runtime nesting depth costs neither registers nor scratch.
// EXT: whether the navigation built on this cursor handles the extended
// geometry - hexagonal lattices and cylinders off the coordinate axes;
// kernels for problems without them are built without that code (the host
// launches the EXT variants for every problem that has one).
template <bool EXT>
struct TrkT {
  static constexpr bool kExt = EXT;
  int32_t np;     // pads in use
  int32_t cell;   // material cell index, -1 = lost
  int32_t token;  // surface token the particle sits on
  int32_t err;    // ABX_E_* raised inside navigation
  int32_t D, L;   // pad and frame capacity
  int32_t st;     // element stride
  int32_t* pad_;  // pad k:        pad_[k * st]
  int32_t* tile_; // tile (k, c):  tile_[(3 k + c) * st]
  double* lv_;    // frame (l, c): lv_[(3 l + c) * st]; frame 0 is global

  ABX_HDM int32_t pad(int k) const { return pad_[k * st]; }
  ABX_HDM int32_t tile(int k, int c) const { return tile_[(3 * k + c) * st]; }
  ABX_HDM V3 lv(int l) const {
    return V3{lv_[(3 * l) * st], lv_[(3 * l + 1) * st], lv_[(3 * l + 2) * st]};
  }
  ABX_HDM void set_lv(int l, const V3& v) {
    lv_[(3 * l) * st] = v.x;
    lv_[(3 * l + 1) * st] = v.y;
    lv_[(3 * l + 2) * st] = v.z;
  }
  ABX_HDM bool tile_is(int k, const int t[3]) const {
    return tile(k, 0) == t[0] && tile(k, 1) == t[1] && tile(k, 2) == t[2];
  }
};

// Bytes of cursor storage per lane for capacities (D, L).
ABX_HD int trk_bytes(int D, int L) { return 16 * D + 24 * L; }

// Point a cursor at a block of `stride` lanes' storage (lane `lane`).
template <class T>
ABX_HD void trk_bind(T& t, void* base, int stride, int lane, int D, int L) {
  int32_t* b = (int32_t*)base;
  t.D = D;
  t.L = L;
  t.st = stride;
  t.pad_ = b + lane;
  t.tile_ = b + (size_t)D * stride + lane;
  t.lv_ = (double*)(b + (size_t)4 * D * stride) + lane;
}

ABX_HD bool tile_in_range(const DUniv& L, const int t[3]) {
  return !(t[0] < 0 || t[0] >= L.Nx || t[1] < 0 || t[1] >= L.Ny || t[2] < 0 ||
           t[2] >= L.Nz);
}

template <class T>
ABX_HD bool trk_push(T& t, int32_t w, const int* tile) {
  if (t.np >= t.D) {
    t.err = -3;
    return false;
  }
  t.pad_[t.np * t.st] = w;
  if (tile) {
    t.tile_[(3 * t.np) * t.st] = tile[0];
    t.tile_[(3 * t.np + 1) * t.st] = tile[1];
    t.tile_[(3 * t.np + 2) * t.st] = tile[2];
  }
  t.np++;
  return true;
}

// Universe::get_cell with the stack, iteratively; r is frame `lvl`.
template <class T>
ABX_NOINL int trk_descend(const DevProblem& P, T& t, int ui, int lvl, V3 r, const V3& u) {
  for (int guard = 0; guard < 4 * ABX_MAXPADS; guard++) {
    const DUniv& U = P.univ[ui];
    if (U.type == ABX_CELL_UNIVERSE) {
      if (!trk_push(t, pad_word(PAD_U, lvl, ui), nullptr)) return -1;
      int found = -1;
      for (int k = 0; k < U.ncells; k++) {
        const int ci = P.ucells[U.cells_off + k];
        if (cell_inside<T::kExt>(P, ci, r, u, t.token)) {
          found = ci;
          break;
        }
      }
      if (found < 0) return -1;
      if (!trk_push(t, pad_word(PAD_C, lvl, found), nullptr)) return -1;
      const DCell& c = P.cell[found];
      if (c.fill_type == 0) return found;
      ui = c.fill;
      continue;
    }
    int tl[3];
    int elem = -1;
    bool hex = false;
    if (T::kExt && U.type == ABX_HEX_LATTICE) {
      // HexLattice::get_cell (src/hex_lattice.cpp:143-217): the tile of the
      // origin-relative point; the pad and the outer universe keep r
      const HexTile h = hex_tile(U, P.elems, r.x - U.Xl, r.y - U.Yl, r.z - U.Zl);
      tl[0] = h.q;
      tl[1] = h.r;
      tl[2] = h.z;
      elem = h.elem;
      hex = true;
    } else {
      get_tile(U, r, u, tl);
      if (tile_in_range(U, tl))
        elem = P.elems[U.elems_off + tl[2] * (U.Nx * U.Ny) + tl[0] * U.Ny + tl[1]];
    }
    if (!trk_push(t, pad_word(PAD_L, lvl, ui), tl)) return -1;
    if (elem >= 0) {
      if (lvl + 1 >= t.L) {
        t.err = -3;
        return -1;
      }
      // the element's frame: (origin-relative point) - tile centre for a
      // hexagonal lattice
      if (T::kExt && hex)
        r = vsub(vsub(r, V3{U.Xl, U.Yl, U.Zl}), hex_tile_center(U, tl[0], tl[1], tl[2]));
      else
        r = vsub(r, tile_center(U, tl[0], tl[1], tl[2]));
      lvl++;
      t.set_lv(lvl, r);
      ui = elem;
      continue;
    }
    if (U.outer >= 0) {
      ui = U.outer;
      continue;
    }
    return -1;
  }
  t.err = -3;
  return -1;
}

// Descend from pad `first_bad` (0 = from the root universe), the lily pads
// below it being stale: Tracker::restart_get_current (first_bad == 0) and
// the partial re-descent of Tracker::get_current (tracker.hpp:235-306),
// which falls back to a restart when the partial descent finds no cell.
// One descent site per call: the fallback loops instead of repeating it.
template <class T>
ABX_HD void trk_locate(const DevProblem& P, T& t, const V3& u, int first_bad) {
  for (;;) {
    int ui, lvl;
    if (first_bad > 0) {
      const int32_t back = t.pad(first_bad - 1);
      lvl = pad_lvl(back);
      ui = pad_type(back) == PAD_C ? P.cell[pad_idx(back)].redescend : pad_idx(back);
      t.np = first_bad - 1;
    } else {
      ui = P.root;
      lvl = 0;
      t.np = 0;
    }
    t.cell = trk_descend(P, t, ui, lvl, t.lv(lvl), u);
    if (t.cell >= 0 || first_bad == 0) return;
    first_bad = 0;
  }
}

// Tracker::restart_get_current from the global position lv[0].
template <class T>
ABX_HD void trk_restart(const DevProblem& P, T& t, const V3& u) { trk_locate(P, t, u, 0); }

template <class T>
ABX_HD void trk_init(const DevProblem& P, T& t, const V3& r, const V3& u, int32_t token) {
  t.err = 0;
  t.token = token;
  t.set_lv(0, r);
  trk_restart(P, t, u);
}

// Tracker::get_current after a move (tree consistent with lv[0]).
template <class T>
ABX_HD void trk_get_current(const DevProblem& P, T& t, const V3& u) {
  int first_bad = t.np;
  for (int i = 0; i < t.np; i++) {
    const int32_t w = t.pad(i);
    const int ty = pad_type(w);
    if (ty == PAD_C) {
      if (!cell_inside<T::kExt>(P, pad_idx(w), t.lv(pad_lvl(w)), u, t.token)) {
        first_bad = i;
        break;
      }
    } else if (ty == PAD_L) {
      int tl[3];
      lat_pad_tile<T::kExt>(P.univ[pad_idx(w)], t.lv(pad_lvl(w)), u, tl);
      if (!t.tile_is(i, tl)) {
        first_bad = i;
        break;
      }
    }
  }
  if (first_bad == t.np) return;
  trk_locate(P, t, u, first_bad);
}

// Tracker::set_r(r) followed by get_current(): the stale tree forces a
// restart, after which the reference still compares against the old tree
// size (tracker.hpp:237-278) and re-descends below it if the new tree is
// deeper.  Both descents go through the one trk_locate site of the loop.
template <class T>
ABX_HD void trk_reset_position(const DevProblem& P, T& t, const V3& r, const V3& u) {
  const int old_np = t.np;
  t.token = 0;
  t.set_lv(0, r);
  int from = 0;
  for (int pass = 0; pass < 2; pass++) {
    trk_locate(P, t, u, from);
    if (pass == 1 || old_np == t.np) return;
    if (old_np > t.np) {
      t.err = -3;  // reference would resize into default pads
      return;
    }
    from = old_np;
  }
}

template <class T>
ABX_HD void trk_move(T& t, const V3& u, double d) {
  const V3 du = vscale(u, d);
  // pads are pushed in descent order, so their frame levels never decrease:
  // the deepest frame in use is the last pad's
  const int maxl = t.np > 0 ? pad_lvl(t.pad(t.np - 1)) : 0;
  for (int l = 0; l <= maxl && l < t.L; l++) t.set_lv(l, vadd(t.lv(l), du));
  t.token = 0;
}

struct Bound {
  double d;
  int32_t sidx;
  int32_t btype;
  int32_t tok;
};

// Universe::get_boundary_condition (cell universe / lattice via outer)
template <bool EXT>
ABX_NOINL Bound univ_bc(const DevProblem& P, int ui, const V3& r, const V3& u, int32_t on) {
  Bound b{ABX_INF, -1, 0, 0};
  for (int guard = 0; guard < 16; guard++) {
    const DUniv& U = P.univ[ui];
    if (EXT ? U.type != ABX_CELL_UNIVERSE : U.type == ABX_RECT_LATTICE) {
      if (!U.has_bc) return b;
      ui = U.outer;
      continue;
    }
    if (U.has_bc) {
      for (int k = 0; k < U.ncells; k++) {
        const int ci = P.ucells[U.cells_off + k];
        if (!P.cell[ci].has_bc) continue;
        const DistTok dt = cell_distance<EXT>(P, ci, r, u, on, true);
        if (dt.d < b.d && dabs(dt.d - b.d) > ABX_BOUNDRY_TOL) {
          const int32_t tt = iabs(dt.tok);
          if (!tt) continue;
          b.tok = tt;
          b.d = dt.d;
          b.sidx = tt - 1;
          b.btype = P.surf[b.sidx].bc;
          if (surf_sign<EXT>(P.surf[b.sidx], r, u) < 0) b.tok = -b.tok;
        }
      }
    }
    return b;
  }
  return b;
}

// Tracker::get_boundary_condition
template <class T>
ABX_NOINL Bound trk_bc(const DevProblem& P, const T& t, const V3& u) {
  if (t.cell < 0) return univ_bc<T::kExt>(P, P.root, t.lv(0), u, t.token);
  Bound b{ABX_INF, -1, 0, 0};
  for (int i = 0; i < t.np; i++) {
    const int32_t w = t.pad(i);
    const V3 rl = t.lv(pad_lvl(w));
    if (pad_type(w) == PAD_C) {
      const int ci = pad_idx(w);
      if (!P.cell[ci].has_bc) continue;
      const DistTok dt = cell_distance<T::kExt>(P, ci, rl, u, t.token, true);
      if (dt.d < b.d && dabs(dt.d - b.d) > ABX_BOUNDRY_TOL) {
        const int32_t tt = iabs(dt.tok);
        if (!tt) continue;
        b.d = dt.d;
        b.tok = tt;
        b.sidx = tt - 1;
        b.btype = P.surf[b.sidx].bc;
        if (surf_sign<T::kExt>(P.surf[b.sidx], rl, u) < 0) b.tok = -b.tok;
      }
    } else {
      const int ui = pad_idx(w);
      if (P.univ[ui].has_bc) {
        const Bound ub = univ_bc<T::kExt>(P, ui, rl, u, t.token);
        if (ub.d < b.d && dabs(ub.d - b.d) > ABX_BOUNDRY_TOL) b = ub;
      }
    }
  }
  return b;
}

// The same out of line, for delta / carter tracking, which only needs it
// when a flight left the geometry.
template <class T>
ABX_HD Bound trk_bc_cold(const DevProblem& P, T t, V3 u) { return trk_bc(P, t, u); }

// Tracker::get_nearest_boundary
template <class T>
ABX_NOINL Bound trk_nearest(const DevProblem& P, const T& t, const V3& u) {
  Bound b = trk_bc(P, t, u);
  for (int i = 0; i < t.np; i++) {
    const int32_t w = t.pad(i);
    const V3 rl = t.lv(pad_lvl(w));
    const int ty = pad_type(w);
    if (ty == PAD_L) {
      const int32_t tl[3] = {t.tile(i, 0), t.tile(i, 1), t.tile(i, 2)};
      const double d = lat_pad_distance<T::kExt>(P.univ[pad_idx(w)], rl, u, tl);
      if (d < b.d && dabs(d - b.d) > ABX_BOUNDRY_TOL) {
        b.d = d;
        b.btype = 2;
        b.sidx = -1;
        b.tok = 0;
      }
    } else if (ty == PAD_C) {
      const DistTok dt = cell_distance<T::kExt>(P, pad_idx(w), rl, u, t.token, false);
      if (dt.d < b.d && dabs(dt.d - b.d) > ABX_BOUNDRY_TOL) {
        b.d = dt.d;
        b.tok = iabs(dt.tok);
        b.sidx = b.tok ? b.tok - 1 : -1;
        b.btype = b.sidx >= 0 ? P.surf[b.sidx].bc : 2;
        if (b.sidx >= 0 && surf_sign<T::kExt>(P.surf[b.sidx], rl, u) < 0) b.tok = -b.tok;
      }
    }
  }
  return b;
}
