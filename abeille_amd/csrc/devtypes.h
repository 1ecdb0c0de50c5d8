// devtypes.h - flattened, HBM-resident problem tables.
//
// Built once by abx_create() from abx_problem_desc and uploaded as one
// allocation; kernels get a DevProblem by value (pointers into that block).
// Layout choices for gfx950:
//  * surfaces are 48-byte records (4 doubles + type/bc) so one 16-byte-aligned
//    burst per surface; cells keep their RPN tokens contiguous;
//  * per-material cross sections are [mat][G] rows, scatter / chi CDFs
//    [mat][G][G] (row = incoming group), small enough (< 40 KB at G = 33 with
//    a handful of materials) to stay resident in L2 / LDS;
//  * mesh tallies are FP64 arrays [Ne][Nx][Ny][Nz] like MeshTally::tally_gen.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define ABX_HD __host__ __device__ __forceinline__
// Bulky device paths.  Inlined by default: an out-of-line call on AMDGPU
// follows the call ABI (callee-saved VGPRs spilled to a scratch frame), which
// cost more than the instruction-cache misses it saved.  -DABX_OUTLINE=1
// builds them out of line.
#if defined(ABX_OUTLINE) && ABX_OUTLINE
#define ABX_NOINL static __host__ __device__ __attribute__((noinline))
#else
#define ABX_NOINL __host__ __device__ __forceinline__
#endif
// Rarely executed paths: always out of line on the device, so that their
// code is not copied into every inlined caller.
#define ABX_COLD static __host__ __device__ __attribute__((noinline))
#else
#define ABX_HD static inline
#define ABX_NOINL static
#define ABX_COLD static
#endif

// Member functions (no `static` on the host).
#if defined(__HIPCC__)
#define ABX_HDM __host__ __device__ __forceinline__
#else
#define ABX_HDM inline
#endif

// Transport kernels: 256-lane blocks, compiled for 3 resident waves per SIMD
// (at most 168 VGPRs; left to itself the compiler takes 176, which allows
// only 2).  The third wave hides more of the LDS and memory latency than the
// few spills it costs: +10-17 % on the lattice, core and batch workloads.
// -DABX_WAVES_PER_EU=n builds another target.
#ifndef ABX_WAVES_PER_EU
#define ABX_WAVES_PER_EU 3
#endif
#define ABX_TKERNEL __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ABX_WAVES_PER_EU)))

#define ABX_MAXPADS 10   // lily pads per tracker (C3 needs 6)
#define ABX_MAXLVL 4     // coordinate frames (root + 3 nested lattices)
#define ABX_SEC_CAP 64   // secondary stack per history slot
#define ABX_INF 1.7976931348623157e308
#define ABX_PI 3.14159265358979323846264338327950288
#define ABX_SURFACE_COINCIDENT 1E-12
#define ABX_BOUNDRY_TOL (500. * 1E-12)

struct DSurf {
  double p[4];
  int32_t type;
  int32_t bc;
  int32_t pad_[2];
};

struct DCell {
  int32_t rpn_off, rpn_len;
  int32_t simple;
  int32_t fill_type, fill;
  int32_t has_bc;
  int32_t redescend;  // universe index used when re-descending from this cell's pad
  uint32_t id;
};

// Universe: cell list, or lattice.  A rectilinear lattice keeps its lower
// corner in Xl/Yl/Zl; a hexagonal one keeps its pitch in Px, axial pitch in
// Pz, origin in Xl/Yl/Zl, tiles in an Nx x Ny (= width x width) x Nz array
// of shifted axial coordinates, and the trigonometric constants of
// include/geometry/hex_lattice.hpp:60-63 as the host computed them.
struct DUniv {
  int32_t type;
  int32_t cells_off, ncells;
  int32_t has_bc;
  int32_t Nx, Ny, Nz;
  int32_t elems_off;
  int32_t outer;
  int32_t top;    // hexagonal: abx_hex_top
  double Px, Py, Pz, Pxi, Pyi, Pzi, Xl, Yl, Zl;
  double z_low, c6, s6, c3, s3;  // hexagonal
  int32_t rings;  // hexagonal: rings; mid = rings - 1
  int32_t pad_;
};

struct DTally {
  int32_t estimator, quantity;
  int32_t Nx, Ny, Nz, Ne;
  int32_t eb_off;   // into DevProblem.tally_eb
  int32_t pad_;
  int64_t data_off; // into tally data array (doubles)
  double lx, ly, lz, hx, hy, hz;
  double dx, dy, dz, dxi, dyi, dzi;
  double net_weight;
};

struct DevProblem {
  const DSurf* surf;
  const DCell* cell;
  const int32_t* rpn;
  const DUniv* univ;
  const int32_t* ucells;
  const int32_t* elems;
  int32_t root;
  int32_t nmat;
  int32_t G;
  int32_t mode;
  int32_t tracking;
  int32_t ntally;
  int32_t n_dgrp_max;
  int32_t precision;
  int32_t trk_D, trk_L;  // cursor capacity (pads, frames) for this geometry
  double wgt_cutoff, wgt_survival, wgt_split;
  double min_energy, max_energy;
  uint64_t seed, stride;
  // energy groups
  const double* ebounds;   // [G+1]
  const double* emid;      // [G] 0.5 * (b_g + b_{g+1})
  const double* majorant;  // [G]
  const double* sampling;  // [G] carter sampling xs
  // per material [nmat][G]
  const double* Et;
  const double* Eabs;    // fission + (Ea - Ef), the micro absorption
  const double* Ef;
  const double* nu_tot;
  const double* nu_d;
  const double* Eel;     // scatter row sums
  const double* vEf;     // nu_total * fission
  const int32_t* fissile;  // [nmat]
  const double* scat_cdf;  // [nmat][G][G]
  const double* chi_cdf;   // [nmat][G][G]
  const int32_t* ang_base; // [nmat] into ang_off, -1 = isotropic material
  const int32_t* ang_off;  // per anisotropic material [G][G] + 1 into ang_mu/pdf/cdf
  const double* ang_mu;
  const double* ang_pdf;
  const double* ang_cdf;
  const int32_t* dgrp_n;   // [nmat] delayed groups
  const double* dgrp_cdf;  // [nmat][n_dgrp_max]
  // tallies
  const DTally* tally;
  const double* tally_eb;
  double* tally_data;
  // per-group tally energy bin of the group midpoint, [ntally][G]
  const int32_t* tally_gbin;
  // Every table above except tally_data lives in one contiguous block
  // [blob, blob + blob_bytes) so that a kernel can stage it in LDS.
  const unsigned char* blob;
  int64_t blob_bytes;
};

// The same problem with its tables read from `base`, a copy of the blob
// (in LDS): every table pointer of dev_view() points into the blob.  The new
// pointers are formed from `base` plus an integer offset, never by adding
// to the old pointer: a pointer member of a kernel argument is assumed to
// address global memory, and arithmetic on it would keep that assumption
// (global loads of LDS addresses fault).
template <class T>
ABX_HD const T* rebase_ptr(const T* p, const unsigned char* blob, const unsigned char* base) {
  const uint64_t off = (uint64_t)((const unsigned char*)p - blob);
  return (const T*)(base + off);
}

ABX_HD DevProblem rebase_problem(const DevProblem& P, const unsigned char* base) {
  DevProblem Q = P;
#define ABX_REBASE(f) Q.f = rebase_ptr(P.f, P.blob, base)
  ABX_REBASE(surf);
  ABX_REBASE(cell);
  ABX_REBASE(rpn);
  ABX_REBASE(univ);
  ABX_REBASE(ucells);
  ABX_REBASE(elems);
  ABX_REBASE(ebounds);
  ABX_REBASE(emid);
  ABX_REBASE(majorant);
  ABX_REBASE(sampling);
  ABX_REBASE(Et);
  ABX_REBASE(Eabs);
  ABX_REBASE(Ef);
  ABX_REBASE(nu_tot);
  ABX_REBASE(nu_d);
  ABX_REBASE(Eel);
  ABX_REBASE(vEf);
  ABX_REBASE(fissile);
  ABX_REBASE(scat_cdf);
  ABX_REBASE(chi_cdf);
  ABX_REBASE(ang_base);
  ABX_REBASE(ang_off);
  ABX_REBASE(ang_mu);
  ABX_REBASE(ang_pdf);
  ABX_REBASE(ang_cdf);
  ABX_REBASE(dgrp_n);
  ABX_REBASE(dgrp_cdf);
  ABX_REBASE(tally);
  ABX_REBASE(tally_eb);
  ABX_REBASE(tally_gbin);
#undef ABX_REBASE
  Q.blob = base;
  return Q;
}
