// kernels.hip -- HIP kernels for gfx950 (MI355X).
//
// Reference passes reproduced (pkestene/euler2d_kokkos src/HydroRunFunctors.h):
//   Init*Functor :1347-1827, MakeBoundariesFunctor :1832-2030, ComputeDtFunctor
//   :17-79, ConvertToPrimitivesFunctor :84-143, ComputeAndStoreFluxesFunctor
//   :412-651, UpdateFunctor :656-723.
// step_fused_kernel performs BC -> prim -> slopes -> trace -> Riemann ->
// update -> next-step dt in one pass over HBM.  Its FP64 instantiation gives
// results bit-identical to the reference pass structure; its split-precision
// instantiation computes in FP32 on a hi+lo FP32 conserved state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

#include "kernels.hpp"

#pragma clang fp contract(off)

namespace e2d {

namespace {

// ----------------------------------------------------- storage policies
// Element (plane, row, column) of a grid addressed as a wave-uniform row base
// plus a 32-bit per-lane byte offset `cb`: the base stays in scalar registers
// and the access uses the scalar-base + vector-offset form of global memory
// instructions, with no per-lane 64-bit address arithmetic.  `row` (and the
// grid) must be the same in every lane of the wavefront: the base is taken
// from the first active lane.
template <class S>
using global_ptr = __attribute__((address_space(1))) S *;

// byte address of the start of `row` in plane 0 (wave-uniform)
template <class S>
__device__ __forceinline__ unsigned long long row_addr(const Grid &g, int row) {
  return reinterpret_cast<unsigned long long>(g.p) +
         (unsigned long long)((long long)row * g.pitch * (long long)sizeof(S));
}

template <class S>
__device__ __forceinline__ global_ptr<S> plane_elem(unsigned long long ra, const Grid &g,
                                                    int plane, unsigned cb) {
  const unsigned long long b =
      ra + (unsigned long long)plane * (unsigned long long)(g.plane * (long long)sizeof(S));
  // readfirstlane keeps the compiler from folding the lane offset into a
  // per-lane 64-bit base (on a value already uniform it emits nothing)
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  global_ptr<char> base = (global_ptr<char>)(((unsigned long long)hi << 32) | lo);
  return (global_ptr<S>)(base + cb);
}

// A loop-invariant lane offset passed through an empty asm once per march
// row: its zero extension then stays next to the row's accesses instead of
// being hoisted out of the loop as a 64-bit value (which loses the
// scalar-base form).  Costs one register move per row.
__device__ __forceinline__ unsigned pin_offset(unsigned cb) {
  asm volatile("" : "+v"(cb));
  return cb;
}

// Element pointers of one row of a grid.  The 64-bit form (O32 = false) puts
// the row's address in scalar registers and works for any grid size.  The O32
// form adds the row's byte offset to the lane's 32-bit offset (one vector add
// per row) and keeps the plane bases loop-invariant, which removes the 64-bit
// scalar address arithmetic of every row; it requires every plane to be at
// most 4 GiB, which the launchers check.
template <class S, bool O32>
struct RowPtr {
  unsigned long long ra;
  unsigned off;
  __device__ __forceinline__ RowPtr(const Grid &g, int row, unsigned cb) {
    if constexpr (O32) {
      ra = reinterpret_cast<unsigned long long>(g.p);
      off = cb + (unsigned)row * ((unsigned)g.pitch * (unsigned)sizeof(S));
    } else {
      ra = row_addr<S>(g, row);
      off = cb;
    }
  }
  __device__ __forceinline__ global_ptr<S> at(const Grid &g, int plane) const {
    return plane_elem<S>(ra, g, plane, off);
  }
};

// FP64: four double planes; the state is the value itself.
struct P64 {
  using T = double;
  static constexpr unsigned kElem = sizeof(double);  // bytes per element of a plane
  struct Cell {
    double u[4];
  };
  static __device__ __forceinline__ Cell load(const Grid &g, long long off) {
    const double *p = static_cast<const double *>(g.p);
    Cell c;
    for (int v = 0; v < 4; v++) c.u[v] = p[v * g.plane + off];
    return c;
  }
  static __device__ __forceinline__ void store(const Grid &g, long long off, const Cell &c) {
    double *p = static_cast<double *>(g.p);
    for (int v = 0; v < 4; v++) p[v * g.plane + off] = c.u[v];
  }
  template <bool O32 = false>
  static __device__ __forceinline__ Cell load_rc(const Grid &g, int row, unsigned cb) {
    const RowPtr<double, O32> rp(g, row, cb);
    Cell c;
    for (int v = 0; v < 4; v++) c.u[v] = *rp.at(g, v);
    return c;
  }
  template <bool O32 = false>
  static __device__ __forceinline__ void store_rc(const Grid &g, int row, unsigned cb,
                                                  const Cell &c) {
    // streaming stores (nt): the rows are read again only by the next
    // launch.  0.8% off the 1024x1536 step; the split policy's stores lose 10%
    // with the same hint and stay plain.
    const RowPtr<double, O32> rp(g, row, cb);
    for (int v = 0; v < 4; v++) __builtin_nontemporal_store(c.u[v], rp.at(g, v));
  }
  static __device__ __forceinline__ void negate(Cell &c, int v) { c.u[v] = -c.u[v]; }
  static __device__ __forceinline__ Cell from_double(double d, double e, double mx, double my) {
    return Cell{{d, e, mx, my}};
  }
  static __device__ __forceinline__ double value(const Cell &c, int v) { return c.u[v]; }
  static __device__ __forceinline__ const SchemeT<T> &scheme(const KParams &kp) { return kp.s; }
  static __device__ __forceinline__ const RateT<T> &rate(const KParams &kp) { return kp.rd; }
  static __device__ __forceinline__ void prim(const KParams &kp, const Cell &c, Q4T<T> &q) {
    prim_no_c(kp.s, c.u[0], c.u[1], c.u[2], c.u[3], q);
  }
  // The march's primitive conversion keeps the reciprocal of the density for
  // the trace's source terms and checks both with one wave-level test.
  static constexpr bool kSharedRecip = true;
  static __device__ __forceinline__ bool prim_shared(const KParams &kp, const Cell &c, Q4T<T> &q,
                                                     double &r, bool &rok) {
    return prim_fp64_shared(kp.s, c.u[0], c.u[1], c.u[2], c.u[3], q, r, rok);
  }
  static __device__ __forceinline__ void prim_ieee(const KParams &kp, const Cell &c, Q4T<T> &q) {
    prim_no_c<double, false>(kp.s, c.u[0], c.u[1], c.u[2], c.u[3], q);
  }
  // internal energy below kin_frac of the total energy (diagnostic only)
  static __device__ __forceinline__ bool kinetic_cell(const KParams &kp, const Cell &c) {
    Q4 q;
    prim(kp, c, q);
    return !(q.v[IP] >= (double)kp.kin_frac * (kp.s.gamma0 - 1.0) * c.u[1]);
  }
  // UpdateFunctor order: (((U + Fx_i) - Fx_{i+1}) + Fy_j) - Fy_{j+1}
  static __device__ __forceinline__ Cell update(const Cell &c, const Q4T<T> &fxl,
                                                const Q4T<T> &fxr, const Q4T<T> &fyb,
                                                const Q4T<T> &fyt) {
    Cell o;
    for (int v = 0; v < 4; v++) {
      double u = c.u[v];
      u += fxl.v[v];
      u -= fxr.v[v];
      u += fyb.v[v];
      u -= fyt.v[v];
      o.u[v] = u;
    }
    return o;
  }
};

// Split precision: one plane per variable whose elements are (hi, lo) FP32
// pairs, about 48 significant bits of state; a lane moves both halves with one
// 8-byte access.  Fluxes are FP32.  The update adds the four face fluxes to
// hi + lo in FP64, in the reference's order, then stores the FP32 rounding of
// the sum as hi and the FP32 rounding of the remainder (exact in FP64) as lo.
// The same FP32 flux value leaves one cell and enters its neighbour, so the
// global sums move only by the roundings of lo, about 2^-48 of a cell's value
// per update.
struct PSplit {
  using T = float;
  static constexpr unsigned kElem = sizeof(float2);
  static constexpr bool kSharedRecip = false;
  struct Cell {
    float hi[4], lo[4];
  };
  static __device__ __forceinline__ Cell load(const Grid &g, long long off) {
    const float2 *p = static_cast<const float2 *>(g.p);
    Cell c;
    for (int v = 0; v < 4; v++) {
      const float2 x = p[v * g.plane + off];
      c.hi[v] = x.x;
      c.lo[v] = x.y;
    }
    return c;
  }
  static __device__ __forceinline__ void store(const Grid &g, long long off, const Cell &c) {
    float2 *p = static_cast<float2 *>(g.p);
    for (int v = 0; v < 4; v++) p[v * g.plane + off] = make_float2(c.hi[v], c.lo[v]);
  }
  // the pair moves as one 64-bit word: low half hi, high half lo
  template <bool O32 = false>
  static __device__ __forceinline__ Cell load_rc(const Grid &g, int row, unsigned cb) {
    const RowPtr<unsigned long long, O32> rp(g, row, cb);
    Cell c;
    for (int v = 0; v < 4; v++) {
      const unsigned long long x = *rp.at(g, v);
      c.hi[v] = __uint_as_float((unsigned)x);
      c.lo[v] = __uint_as_float((unsigned)(x >> 32));
    }
    return c;
  }
  template <bool O32 = false>
  static __device__ __forceinline__ void store_rc(const Grid &g, int row, unsigned cb,
                                                  const Cell &c) {
    const RowPtr<unsigned long long, O32> rp(g, row, cb);
    for (int v = 0; v < 4; v++)
      *rp.at(g, v) = ((unsigned long long)__float_as_uint(c.lo[v]) << 32) |
                     __float_as_uint(c.hi[v]);
  }
  static __device__ __forceinline__ void negate(Cell &c, int v) {
    c.hi[v] = -c.hi[v];
    c.lo[v] = -c.lo[v];
  }
  static __device__ __forceinline__ Cell from_double(double d, double e, double mx, double my) {
    const double x[4] = {d, e, mx, my};
    Cell c;
    for (int v = 0; v < 4; v++) {
      c.hi[v] = (float)x[v];
      c.lo[v] = (float)(x[v] - (double)c.hi[v]);
    }
    return c;
  }
  static __device__ __forceinline__ double value(const Cell &c, int v) {
    return (double)c.hi[v] + (double)c.lo[v];
  }
  static __device__ __forceinline__ const SchemeT<T> &scheme(const KParams &kp) { return kp.sf; }
  static __device__ __forceinline__ const RateT<T> &rate(const KParams &kp) { return kp.rf; }
  // FP32 conversion; cells whose internal energy is a small fraction of the
  // total energy (kinetic-dominated: p = (gamma-1)(E - rho v^2/2) cancels)
  // are converted again in FP64 from hi + lo.
  static __device__ __forceinline__ void prim32(const KParams &kp, const Cell &c, Q4T<T> &q) {
    prim_no_c(kp.sf, c.hi[0] + c.lo[0], c.hi[1] + c.lo[1], c.hi[2] + c.lo[2], c.hi[3] + c.lo[3],
              q);
  }
  // the FP32 pressure below kin_frac (gamma-1) E: internal energy below that
  // fraction of the total (false for every cell when kin_frac = 0)
  static __device__ __forceinline__ bool kinetic(const KParams &kp, const Cell &c,
                                                 const Q4T<T> &q) {
    const float etot = c.hi[1];
    return !(q.v[IP] >= kp.kin_frac * (kp.sf.gamma0 - 1.0f) * etot);
  }
  static __device__ __forceinline__ void prim(const KParams &kp, const Cell &c, Q4T<T> &q) {
    prim32(kp, c, q);
    if (kinetic(kp, c, q)) {
      Q4 q64;
      prim_no_c<double, false>(kp.s, value(c, 0), value(c, 1), value(c, 2), value(c, 3), q64);
      for (int v = 0; v < 4; v++) q.v[v] = (float)q64.v[v];
    }
  }
  // the cell takes the FP64 conversion
  static __device__ __forceinline__ bool kinetic_cell(const KParams &kp, const Cell &c) {
    Q4T<T> q;
    prim32(kp, c, q);
    return kinetic(kp, c, q);
  }
  static __device__ __forceinline__ Cell update(const Cell &c, const Q4T<T> &fxl,
                                                const Q4T<T> &fxr, const Q4T<T> &fyb,
                                                const Q4T<T> &fyt) {
    Cell o;
    for (int v = 0; v < 4; v++) {
      double x = (double)c.hi[v] + (double)c.lo[v];
      x += (double)fxl.v[v];
      x -= (double)fxr.v[v];
      x += (double)fyb.v[v];
      x -= (double)fyt.v[v];
      const float hi = (float)x;
      o.hi[v] = hi;
      o.lo[v] = (float)(x - (double)hi);
    }
    return o;
  }
};

// Plain FP32 (the reference built with USE_DOUBLE=OFF, CMakeLists.txt:48-51):
// four float planes, float arithmetic, the update in float.  A comparator for
// the split path, not a drop-in.
struct P32 {
  using T = float;
  static constexpr unsigned kElem = sizeof(float);
  static constexpr bool kSharedRecip = false;
  struct Cell {
    float u[4];
  };
  static __device__ __forceinline__ Cell load(const Grid &g, long long off) {
    const float *p = static_cast<const float *>(g.p);
    Cell c;
    for (int v = 0; v < 4; v++) c.u[v] = p[v * g.plane + off];
    return c;
  }
  static __device__ __forceinline__ void store(const Grid &g, long long off, const Cell &c) {
    float *p = static_cast<float *>(g.p);
    for (int v = 0; v < 4; v++) p[v * g.plane + off] = c.u[v];
  }
  template <bool O32 = false>
  static __device__ __forceinline__ Cell load_rc(const Grid &g, int row, unsigned cb) {
    const RowPtr<float, O32> rp(g, row, cb);
    Cell c;
    for (int v = 0; v < 4; v++) c.u[v] = *rp.at(g, v);
    return c;
  }
  template <bool O32 = false>
  static __device__ __forceinline__ void store_rc(const Grid &g, int row, unsigned cb,
                                                  const Cell &c) {
    const RowPtr<float, O32> rp(g, row, cb);
    for (int v = 0; v < 4; v++) *rp.at(g, v) = c.u[v];
  }
  static __device__ __forceinline__ void negate(Cell &c, int v) { c.u[v] = -c.u[v]; }
  static __device__ __forceinline__ Cell from_double(double d, double e, double mx, double my) {
    return Cell{{(float)d, (float)e, (float)mx, (float)my}};
  }
  static __device__ __forceinline__ double value(const Cell &c, int v) { return (double)c.u[v]; }
  static __device__ __forceinline__ const SchemeT<T> &scheme(const KParams &kp) { return kp.sf; }
  static __device__ __forceinline__ const RateT<T> &rate(const KParams &kp) { return kp.rf; }
  static __device__ __forceinline__ void prim(const KParams &kp, const Cell &c, Q4T<T> &q) {
    prim_no_c(kp.sf, c.u[0], c.u[1], c.u[2], c.u[3], q);
  }
  static __device__ __forceinline__ bool kinetic_cell(const KParams &kp, const Cell &c) {
    Q4T<T> q;
    prim(kp, c, q);
    return !(q.v[IP] >= kp.kin_frac * (kp.sf.gamma0 - 1.0f) * c.u[1]);
  }
  static __device__ __forceinline__ Cell update(const Cell &c, const Q4T<T> &fxl,
                                                const Q4T<T> &fxr, const Q4T<T> &fyb,
                                                const Q4T<T> &fyt) {
    Cell o;
    for (int v = 0; v < 4; v++) {
      float u = c.u[v];
      u += fxl.v[v];
      u -= fxr.v[v];
      u += fyb.v[v];
      u -= fyt.v[v];
      o.u[v] = u;
    }
    return o;
  }
};

// CFL rate of one cell (ComputeDtFunctor body); FS: the FP64 sound speed
// takes the short square root (kernels built for the host's floor_sqrt);
// POW2: the cell sizes are powers of two
template <class P, bool FS = false, bool POW2 = false>
__device__ __forceinline__ typename P::T cell_rate(const KParams &kp, const typename P::Cell &c) {
  if constexpr (P::kSharedRecip) {
    return cfl_rate_fp64<FS, POW2>(kp.s, c.u[0], c.u[1], c.u[2], c.u[3], kp.rd);
  } else {
    Q4T<typename P::T> q;
    P::prim(kp, c, q);
    return cfl_rate_q(P::scheme(kp), q, P::rate(kp));
  }
}

__device__ __forceinline__ double &at(const Grid &g, int v, int i, int j) {
  return static_cast<double *>(g.p)[v * g.plane + (long long)j * g.pitch + i];
}

// Ghost-index map of MakeBoundariesFunctor (HydroRunFunctors.h:1876-2024) in
// closed form: returns the interior index a ghost index copies from, and
// whether the normal velocity flips sign (BC 1, reflecting).  BC 2 copies the
// first/last interior cell, every other value is periodic.  The reference
// fills x faces then y faces, so a corner ghost is the y rule applied to the x
// rule: both maps compose, and the sign flips (exact) commute.
__device__ __forceinline__ int ghost_map(int i, int n, int bclo, int bchi, bool &flip) {
  // for a row index this is uniform control flow on the scalar unit
  flip = false;
  if (i < 2) {
    if (bclo == 1) {
      flip = true;
      return 3 - i;
    }
    if (bclo == 2) return 2;
    return n + i;
  }
  if (i > n + 1) {
    if (bchi == 1) {
      flip = true;
      return 2 * n + 3 - i;
    }
    if (bchi == 2) return n + 1;
    return i - n;
  }
  return i;
}

// The same map for a wave-uniform row index, written as selects so that it
// compiles to scalar selects rather than a tree of scalar branches.
__device__ __forceinline__ int ghost_map_uniform(int i, int n, int bclo, int bchi, bool &flip) {
  const int lo = bclo == 1 ? 3 - i : (bclo == 2 ? 2 : n + i);
  const int hi = bchi == 1 ? 2 * n + 3 - i : (bchi == 2 ? n + 1 : i - n);
  const bool below = i < 2, above = i > n + 1;
  flip = (below && bclo == 1) || (above && bchi == 1);
  return below ? lo : (above ? hi : i);
}

__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

__device__ __forceinline__ double wave_max(double x) {
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off));
  return x;
}

__device__ __forceinline__ float wave_max(float x) {
  for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off));
  return x;
}

// ------------------------------------------------------------------ init
// Init*Functor bodies in FP64 (HydroRunFunctors.h:1347-1827); the split path
// stores the FP64 value as hi + lo.
template <class P>
__global__ void init_kernel(e2d_params p, Grid U, unsigned int *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= p.isize || j >= p.jsize) return;
  const int gw = p.ghostWidth;
  const double gamma0 = p.gamma0;
  const double x = p.xmin + p.dx / 2 + (i - gw) * p.dx;
  const double y = p.ymin + p.dy / 2 + (j - gw) * p.dy;
  double d, e, mx = 0.0, my = 0.0;
  if (p.problemType == 1) {
    // InitBlastFunctor, TagRegularInit (:1469-1505)
    const double radius2 = p.blast_radius * p.blast_radius;
    const double d2 = (x - p.blast_center_x) * (x - p.blast_center_x) +
                      (y - p.blast_center_y) * (y - p.blast_center_y);
    if (d2 < radius2) {
      d = p.blast_density_in;
      e = p.blast_pressure_in / (gamma0 - 1.0);
      if (count) atomicAdd(count, 1u);
    } else {
      d = p.blast_density_out;
      e = p.blast_pressure_out / (gamma0 - 1.0);
    }
  } else if (p.problemType == 2) {
    // InitFourQuadrantFunctor (:1549-1641), primToCons :1578-1591
    const double W[4][4] = {{1.5, 1.5, 0.0, 0.0},
                            {0.5323, 0.3, 1.206, 0.0},
                            {0.138, 0.029, 1.206, 1.206},
                            {0.5323, 0.3, 0.0, 1.206}};
    const int k = (x < 0.8) ? ((y < 0.8) ? 2 : 1) : ((y < 0.8) ? 3 : 0);
    const double rho = W[k][0], pr = W[k][1], u = W[k][2], v = W[k][3];
    d = rho;
    mx = u * rho;
    my = v * rho;
    e = pr / (gamma0 - 1.0) + rho * (u * u + v * v) * 0.5;
  } else if (p.problemType == 3) {
    // InitDiscontinuityFunctor (:1673-1707)
    if (x + y < 1) {
      d = 1.0 + x * x;
      e = 1.0 / (gamma0 - 1.0);
    } else {
      d = 0.25;
      e = 1.0 / (gamma0 - 1.0);
    }
  } else if (p.problemType == 4) {
    // InitShockedBubbleFunctor (:1741-1820)
    if (x < p.sb_shock_loc) {
      d = p.sb_postshock_density;
      mx = p.sb_postshock_density * p.sb_postshock_velocity;
      const double rho_eint = p.sb_postshock_pressure / (gamma0 - 1);
      e = rho_eint + 0.5 * (mx * mx + my * my) / d;
    } else {
      const double radius = sqrt((x - p.sb_bubble_center_x) * (x - p.sb_bubble_center_x) +
                                 (y - p.sb_bubble_center_y) * (y - p.sb_bubble_center_y));
      if (radius < p.sb_bubble_radius) {
        d = p.sb_bubble_density;
        const double rho_eint = p.sb_bubble_pressure / (gamma0 - 1);
        e = rho_eint + 0.5 * (mx * mx + my * my) / d;
      } else {
        d = p.sb_preshock_density;
        const double rho_eint = p.sb_preshock_pressure / (gamma0 - 1);
        e = rho_eint + 0.5 * (mx * mx + my * my) / d;
      }
    }
  } else {
    // InitImplodeFunctor (:1371-1397); high-state pressure 1 unless overridden
    const double tmp = x + y * y;
    if (tmp > 0.5 && tmp < 1.5) {
      d = 1.0;
      e = p.implode_pressure_in / (gamma0 - 1.0);
    } else {
      d = 0.125;
      e = 0.14 / (gamma0 - 1.0);
    }
  }
  P::store(U, (long long)j * U.pitch + i, P::from_double(d, e, mx, my));
}

// InitBlastFunctor, TagCorrectTotalEnergy (:1507-1523)
template <class P>
__global__ void blast_energy_kernel(e2d_params p, Grid U, double e_inside) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= p.isize || j >= p.jsize) return;
  const double x = p.xmin + p.dx / 2 + (i - p.ghostWidth) * p.dx;
  const double y = p.ymin + p.dy / 2 + (j - p.ghostWidth) * p.dy;
  const double radius2 = p.blast_radius * p.blast_radius;
  const double d2 = (x - p.blast_center_x) * (x - p.blast_center_x) +
                    (y - p.blast_center_y) * (y - p.blast_center_y);
  if (d2 < radius2) {
    const long long off = (long long)j * U.pitch + i;
    typename P::Cell c = P::load(U, off);
    P::store(U, off, P::from_double(P::value(c, 0), e_inside, P::value(c, 2), P::value(c, 3)));
  }
}

// --------------------------------------------------------------- ghosts
// One thread per ghost cell: 2 rows at the bottom and top (all columns), then
// 2 columns on each side of every interior row.
template <class P>
__global__ void boundaries_kernel(KParams kp, Grid U) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int rows = 4 * kp.isize;
  int i, j;
  if (t < rows) {
    const int r = t / kp.isize;
    i = t - r * kp.isize;
    j = r < 2 ? r : kp.ny + r;  // rows 0,1,ny+2,ny+3
  } else {
    const int t2 = t - rows;
    if (t2 >= 4 * kp.ny) return;
    j = 2 + (t2 >> 2);
    const int c = t2 & 3;
    i = c < 2 ? c : kp.nx + c;  // columns 0,1,nx+2,nx+3
  }
  bool fx, fy;
  const int i0 = ghost_map(i, kp.nx, kp.bcxmin, kp.bcxmax, fx);
  const int j0 = ghost_map(j, kp.ny, kp.bcymin, kp.bcymax, fy);
  typename P::Cell c = P::load(U, (long long)j0 * U.pitch + i0);
  if (fx) P::negate(c, IU);
  if (fy) P::negate(c, IV);
  P::store(U, (long long)j * U.pitch + i, c);
}

// ------------------------------------------------------------------- dt
template <class P>
__global__ __launch_bounds__(256) void dt_reduce_kernel(KParams kp, Grid U,
                                                        unsigned long long *invdt_bits) {
  using T = typename P::T;
  T m = T(0.0);
  // block b: interior rows 2+b, 2+b+G, ...; threads stride along the row
  for (int j = 2 + blockIdx.x; j <= kp.ny + 1; j += gridDim.x)
    for (int i = 2 + threadIdx.x; i <= kp.nx + 1; i += blockDim.x)
      m = Ops<T>::mx(m, cell_rate<P>(kp, P::load(U, (long long)j * U.pitch + i)));
  // block max, then one device-scope atomic per block (the max is exact and
  // order independent)
  __shared__ double wm[256 / kLanes];
  const double md = (double)wave_max(m);
  if ((threadIdx.x & (kLanes - 1)) == 0) wm[threadIdx.x / kLanes] = md;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kLanes); w++)
      if (wm[w] == wm[w] && wm[w] > b) b = wm[w];
    if (b > 0.0) atomicMax(invdt_bits, dbits(b));
  }
}

// The copy of a ring entry in host memory (when the control block has one):
// system-scope stores reach host memory past the device caches, so the host
// sees them once the launch has completed.
__device__ __forceinline__ void mirror_state(StepState *hs, int ph, const StepState &n) {
  if (!hs) return;
  unsigned long long *w = reinterpret_cast<unsigned long long *>(hs + ph);
  __hip_atomic_store(w, (unsigned long long)__double_as_longlong(n.t), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(w + 1, (unsigned long long)(unsigned)n.nstep, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

// Start of a run loop in one launch (main.cpp:90-91 and the first compute_dt,
// HydroRun.h:227-251): block b folds the CFL rate of interior rows b, b + G,
// ... into partials[b]; with `bc` the grid's threads also fill the ghost cells
// of `in` and `other` (make_boundaries of both arrays before step 0: the ghost
// cells written are disjoint from the interior cells read).  The last block
// to arrive (atomicInc wraps the arrival counter back to 0 for the next use)
// takes the max of the partials and sets ring phase 0: the state before the
// step and its rate in s[0] / rate[0], rate[1] cleared for launch 0's fold.
// Launch 0 clears rate[2] itself.  Blocks of 16 wavefronts keep the loads of
// many rows in flight (with 4 the kernel waited on load latency).
constexpr int kBeginThreads = 1024;

template <class P>
__global__ __launch_bounds__(kBeginThreads) void loop_begin_kernel(KParams kp, Grid in,
                                                                   Grid other, int bc,
                                                                   StepCtl *ctl, double t,
                                                                   int nstep, double *partials,
                                                                   unsigned *arrivals) {
  using T = typename P::T;
  if (bc) {  // one thread per ghost cell of each array, as boundaries_kernel
    const int nghost = 4 * kp.isize + 4 * kp.ny;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < 2 * nghost;
         g += gridDim.x * blockDim.x) {
      const Grid &U = g < nghost ? in : other;
      const int t0 = g < nghost ? g : g - nghost;
      int i, j;
      if (t0 < 4 * kp.isize) {
        const int r = t0 / kp.isize;
        i = t0 - r * kp.isize;
        j = r < 2 ? r : kp.ny + r;
      } else {
        const int t2 = t0 - 4 * kp.isize;
        j = 2 + (t2 >> 2);
        const int c = t2 & 3;
        i = c < 2 ? c : kp.nx + c;
      }
      bool fx, fy;
      const int i0 = ghost_map(i, kp.nx, kp.bcxmin, kp.bcxmax, fx);
      const int j0 = ghost_map(j, kp.ny, kp.bcymin, kp.bcymax, fy);
      typename P::Cell c = P::load(U, (long long)j0 * U.pitch + i0);
      if (fx) P::negate(c, IU);
      if (fy) P::negate(c, IV);
      P::store(U, (long long)j * U.pitch + i, c);
    }
  }
  // rows b, b + G, b + 2G, b + 3G loaded together, then converted
  constexpr int kRows = 4;
  T m = T(0.0);
  for (int j0 = 2 + blockIdx.x; j0 <= kp.ny + 1; j0 += kRows * gridDim.x)
    for (int i = 2 + threadIdx.x; i <= kp.nx + 1; i += blockDim.x) {
      typename P::Cell c[kRows];
#pragma unroll
      for (int u = 0; u < kRows; u++) {
        const int j = j0 + u * (int)gridDim.x;
        if (j <= kp.ny + 1) c[u] = P::load(in, (long long)j * in.pitch + i);
      }
#pragma unroll
      for (int u = 0; u < kRows; u++) {
        if (j0 + u * (int)gridDim.x <= kp.ny + 1) m = Ops<T>::mx(m, cell_rate<P>(kp, c[u]));
      }
    }
  __shared__ double wm[kBeginThreads / kLanes];
  __shared__ bool last;
  const double md = (double)wave_max(m);
  if ((threadIdx.x & (kLanes - 1)) == 0) wm[threadIdx.x / kLanes] = md;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kLanes); w++)
      if (wm[w] == wm[w] && wm[w] > b) b = wm[w];
    // The partial is written at device scope (write-through, no L2 write-back
    // of the block's other lines) and has completed before the arrival count;
    // the last block reads the partials at device scope.  A device-scope fence
    // in every block instead (an L2 write-back each) made the launch 38 us
    // against 21 at 1024x1536.
    __hip_atomic_store(&partials[blockIdx.x], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    last = atomicInc(arrivals, gridDim.x - 1) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  // the last block: max over the partials (every block's store is visible)
  double b = 0.0;
  for (int k = threadIdx.x; k < (int)gridDim.x; k += blockDim.x) {
    const double x = __hip_atomic_load(&partials[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b = x > b ? x : b;
  }
  b = wave_max(b);
  if ((threadIdx.x & (kLanes - 1)) == 0) wm[threadIdx.x / kLanes] = b;
  __syncthreads();
  if (threadIdx.x < kRateSlots) {
    double mx = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kLanes); w++) mx = wm[w] > mx ? wm[w] : mx;
    ctl->rate[0][threadIdx.x * kSlotStride] = threadIdx.x == 0 ? dbits(mx) : 0ull;
    ctl->rate[1][threadIdx.x * kSlotStride] = 0ull;
  }
  if (threadIdx.x == 0) {
    StepState s0;
    s0.t = t;
    s0.nstep = nstep;
    s0.pad = 0;
    ctl->s[0] = s0;
    mirror_state(ctl->host_s, 0, s0);
  }
}

// Interior cells whose internal energy is below kin_frac of the total energy:
// for the split precision exactly the cells whose primitive variables the step
// converts again in FP64.  One atomic per wavefront.
template <class P>
__global__ __launch_bounds__(256) void kinetic_count_kernel(KParams kp, Grid U,
                                                            unsigned long long *count) {
  unsigned n = 0;
  for (int j = 2 + blockIdx.x; j <= kp.ny + 1; j += gridDim.x)
    for (int i = 2 + threadIdx.x; i <= kp.nx + 1; i += blockDim.x)
      n += P::kinetic_cell(kp, P::load(U, (long long)j * U.pitch + i)) ? 1u : 0u;
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
  if ((threadIdx.x & (kLanes - 1)) == 0 && n) atomicAdd(count, (unsigned long long)n);
}

// ------------------------------------------------ reference pass schedule
__global__ void prim_kernel(KParams kp, Grid U, Grid Q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= kp.isize || j >= kp.jsize) return;
  Q4 q;
  prim_no_c(kp.s, at(U, ID, i, j), at(U, IP, i, j), at(U, IU, i, j), at(U, IV, i, j), q);
  for (int v = 0; v < 4; v++) at(Q, v, i, j) = q.v[v];
}

__device__ __forceinline__ Q4 ld4(const Grid &g, int i, int j) {
  Q4 q;
  for (int v = 0; v < 4; v++) q.v[v] = at(g, v, i, j);
  return q;
}

__device__ __forceinline__ void slopes(const Scheme &s, const Q4 &q, const Q4 &px, const Q4 &mx,
                                       const Q4 &py, const Q4 &my, Q4 &dqx, Q4 &dqy) {
  for (int v = 0; v < 4; v++) {
    dqx.v[v] = slope1(s, q.v[v], px.v[v], mx.v[v]);
    dqy.v[v] = slope1(s, q.v[v], py.v[v], my.v[v]);
  }
}

// ComputeAndStoreFluxesFunctor (:449-644), i,j in [2, size-2]
template <int RS>
__global__ void fluxes_kernel(KParams kp, Grid Q, Grid Fx, Grid Fy, double dtdx, double dtdy) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > kp.isize - 2 || j > kp.jsize - 2) return;
  const Scheme &s = kp.s;
  const Q4 qc = ld4(Q, i, j);
  Q4 dqX, dqY, dqXn, dqYn, fr, fl;
  slopes(s, qc, ld4(Q, i + 1, j), ld4(Q, i - 1, j), ld4(Q, i, j + 1), ld4(Q, i, j - 1), dqX, dqY);
  const Sources sc = trace_sources(s, qc, dqX, dqY);
  double flux[4];
  {
    const Q4 qn = ld4(Q, i - 1, j);
    slopes(s, qn, qc, ld4(Q, i - 2, j), ld4(Q, i - 1, j + 1), ld4(Q, i - 1, j - 1), dqXn, dqYn);
    const Sources sn = trace_sources(s, qn, dqXn, dqYn);
    trace_face(s, qc, dqX, sc, dtdx, false, fr);
    trace_face(s, qn, dqXn, sn, dtdx, true, fl);
    riemann<RS>(s, fl, fr, IU, IV, flux);
    for (int v = 0; v < 4; v++) at(Fx, v, i, j) = flux[v] * dtdx;
  }
  {
    const Q4 qn = ld4(Q, i, j - 1);
    slopes(s, qn, ld4(Q, i + 1, j - 1), ld4(Q, i - 1, j - 1), qc, ld4(Q, i, j - 2), dqXn, dqYn);
    const Sources sn = trace_sources(s, qn, dqXn, dqYn);
    trace_face(s, qc, dqY, sc, dtdy, false, fr);
    trace_face(s, qn, dqYn, sn, dtdy, true, fl);
    riemann<RS>(s, fl, fr, IV, IU, flux);
    for (int v = 0; v < 4; v++) at(Fy, v, i, j) = flux[v] * dtdy;
  }
}

// UpdateFunctor (:683-719)
__global__ void update_kernel(KParams kp, Grid U, Grid Fx, Grid Fy) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > kp.nx + 1 || j > kp.ny + 1) return;
  for (int v = 0; v < 4; v++) {
    double u = at(U, v, i, j);
    u += at(Fx, v, i, j);
    u -= at(Fx, v, i + 1, j);
    u += at(Fy, v, i, j);
    u -= at(Fy, v, i, j + 1);
    at(U, v, i, j) = u;
  }
}

// ---------------------------------- implementation version 1 (directional)
// ComputeSlopesFunctor (HydroRunFunctors.h:1061-1162): slopes stored once,
// for i, j in [1, size-2]
__global__ void slopes_kernel(KParams kp, Grid Q, Grid Sx, Grid Sy) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > kp.isize - 2 || j > kp.jsize - 2) return;
  Q4 dqX, dqY;
  slopes(kp.s, ld4(Q, i, j), ld4(Q, i + 1, j), ld4(Q, i - 1, j), ld4(Q, i, j + 1),
         ld4(Q, i, j - 1), dqX, dqY);
  for (int v = 0; v < 4; v++) {
    at(Sx, v, i, j) = dqX.v[v];
    at(Sy, v, i, j) = dqY.v[v];
  }
}

// ComputeTraceAndFluxes_Functor<dir> (:1167-1342): the flux through the
// left (DIR 0) or bottom (DIR 1) face of (i, j), i, j in [2, size-2]
template <int RS, int DIR>
__global__ void trace_fluxes_kernel(KParams kp, Grid Q, Grid Sx, Grid Sy, Grid F, double dtdx,
                                    double dtdy) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > kp.isize - 2 || j > kp.jsize - 2) return;
  const Scheme &s = kp.s;
  const int in = DIR == 0 ? i - 1 : i, jn = DIR == 0 ? j : j - 1;
  const Q4 qc = ld4(Q, i, j), qn = ld4(Q, in, jn);
  const Q4 dqX = ld4(Sx, i, j), dqY = ld4(Sy, i, j);
  const Q4 dqXn = ld4(Sx, in, jn), dqYn = ld4(Sy, in, jn);
  const Sources sc = trace_sources(s, qc, dqX, dqY);
  const Sources sn = trace_sources(s, qn, dqXn, dqYn);
  const double dtd = DIR == 0 ? dtdx : dtdy;
  Q4 fr, fl;
  trace_face(s, qc, DIR == 0 ? dqX : dqY, sc, dtd, false, fr);
  trace_face(s, qn, DIR == 0 ? dqXn : dqYn, sn, dtd, true, fl);
  double flux[4];
  riemann<RS>(s, fl, fr, DIR == 0 ? IU : IV, DIR == 0 ? IV : IU, flux);
  for (int v = 0; v < 4; v++) at(F, v, i, j) = flux[v] * dtd;
}

// UpdateDirFunctor<dir> (:986-1055): U += F(face before); U -= F(face after)
template <int DIR>
__global__ void update_dir_kernel(KParams kp, Grid U, Grid F) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > kp.nx + 1 || j > kp.ny + 1) return;
  for (int v = 0; v < 4; v++) {
    double u = at(U, v, i, j);
    u += at(F, v, i, j);
    u -= DIR == 0 ? at(F, v, i + 1, j) : at(F, v, i, j + 1);
    at(U, v, i, j) = u;
  }
}

// ---------------------------------- implementation version 2 (atomics)
// ComputeFluxesAndUpdateFunctor (:728-980): the left and bottom face fluxes of
// (i, j), i, j in [2, size-2], added to the cells on both sides with FP64
// atomics.  The order of the additions into a cell is not fixed, so the last
// bits of the result vary from run to run, as in the reference.
template <int RS>
__global__ void fluxes_atomic_kernel(KParams kp, Grid Q, Grid U, double dtdx, double dtdy) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > kp.isize - 2 || j > kp.jsize - 2) return;
  const Scheme &s = kp.s;
  const Q4 qc = ld4(Q, i, j);
  Q4 dqX, dqY, dqXn, dqYn, fr, fl;
  slopes(s, qc, ld4(Q, i + 1, j), ld4(Q, i - 1, j), ld4(Q, i, j + 1), ld4(Q, i, j - 1), dqX, dqY);
  const Sources sc = trace_sources(s, qc, dqX, dqY);
  double flux[4];
  const bool inner_i = i < kp.isize - 2, inner_j = j < kp.jsize - 2;
  {
    const Q4 qn = ld4(Q, i - 1, j);
    slopes(s, qn, qc, ld4(Q, i - 2, j), ld4(Q, i - 1, j + 1), ld4(Q, i - 1, j - 1), dqXn, dqYn);
    const Sources sn = trace_sources(s, qn, dqXn, dqYn);
    trace_face(s, qc, dqX, sc, dtdx, false, fr);
    trace_face(s, qn, dqXn, sn, dtdx, true, fl);
    riemann<RS>(s, fl, fr, IU, IV, flux);
    for (int v = 0; v < 4; v++) {
      const double f = flux[v] * dtdx;
      if (inner_j && inner_i) unsafeAtomicAdd(&at(U, v, i, j), f);
      if (inner_j && i > 2) unsafeAtomicAdd(&at(U, v, i - 1, j), -f);
    }
  }
  {
    const Q4 qn = ld4(Q, i, j - 1);
    slopes(s, qn, ld4(Q, i + 1, j - 1), ld4(Q, i - 1, j - 1), qc, ld4(Q, i, j - 2), dqXn, dqYn);
    const Sources sn = trace_sources(s, qn, dqXn, dqYn);
    trace_face(s, qc, dqY, sc, dtdy, false, fr);
    trace_face(s, qn, dqYn, sn, dtdy, true, fl);
    riemann<RS>(s, fl, fr, IV, IU, flux);
    for (int v = 0; v < 4; v++) {
      const double f = flux[v] * dtdy;
      if (inner_j && inner_i) unsafeAtomicAdd(&at(U, v, i, j), f);
      if (j > 2 && inner_i) unsafeAtomicAdd(&at(U, v, i, j - 1), -f);
    }
  }
}

// ------------------------------------------------------ fused step kernel
//
// One wavefront owns a strip of 64 columns (60 output columns + 2 halo columns
// per side) and marches up `seg_rows` output rows.  Per loaded row r it
// converts U to primitives once, then for row a = r-1 computes the slopes
// once, the Hancock source terms once and the face states it needs, the x-face
// fluxes of row a and the y-face flux between rows a-1 and a, and finally
// updates row b = r-2 with the reference's operation order
// (((U + Fx_i) - Fx_{i+1}) + Fy_j) - Fy_{j+1}.  Horizontal neighbours come from
// lane shuffles, vertical ones from registers carried between iterations, so
// nothing is staged through LDS and every cell is read from HBM once (plus
// the 4-row / 4-column halo overlap between neighbouring waves).
// Ghost cells are produced on the fly from the closed-form BC map; the wave
// owning a ghost cell also writes it to `in` and `out`, as make_boundaries(in)
// followed by deep_copy(out, in) leaves them in the reference.
// In run-loop mode (ctl != nullptr) every wavefront derives t, dt and the
// while condition from the control rings, one wavefront publishes the state
// after the step, and the CFL rate of every updated cell is folded into a
// wave max for the next launch's dt.

// Whole-wavefront lane shifts by one with DPP (GFX9 wave_shr:1 / wave_shl:1):
// a full-rate VALU move instead of an LDS-routed permute.  Lane 0 (shift up)
// and lane 63 (shift down) read zero (bound_ctrl); the march never uses the
// results of those lanes: the x-face of lane 1, the slopes of lanes 0 and 63
// and the x-face of lane 63 feed only halo lanes.
constexpr int kDppWaveShr1 = 0x138;  // lane i reads lane i-1
constexpr int kDppWaveShl1 = 0x130;  // lane i reads lane i+1

template <int CTRL>
__device__ __forceinline__ float dpp_shift(float x) {
  const int v = __float_as_int(x);
  return __int_as_float(__builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true));
}

template <int CTRL>
__device__ __forceinline__ double dpp_shift(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
  const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)rhi << 32) | (unsigned int)rlo);
}

template <class T>
__device__ __forceinline__ Q4T<T> shfl_up1(const Q4T<T> &a) {
  Q4T<T> r;
  for (int v = 0; v < 4; v++) r.v[v] = dpp_shift<kDppWaveShr1>(a.v[v]);
  return r;
}

template <class T>
__device__ __forceinline__ Q4T<T> shfl_dn1(const Q4T<T> &a) {
  Q4T<T> r;
  for (int v = 0; v < 4; v++) r.v[v] = dpp_shift<kDppWaveShl1>(a.v[v]);
  return r;
}

// Stages of one march iteration over loaded row r (row a = r-1, b = r-2).
enum : int {
  kSlopes = 1,  // slopes, source terms and y-face states of row a
  kYFace = 2,   // y-face flux between rows a-1 and a
  kXFace = 4,   // x-face fluxes of row a
  kUpdate = 8,  // update of row b (+ its next-step CFL rate)
  kAll = 15
};

// PD: rows a wavefront loads ahead of the row it converts, 1 (row r + 1 is in
// flight while row r is processed) or 2 (rows r + 1 and r + 2), which hides
// more memory latency at the cost of one more row of registers.  Measured on
// 60-step runs: distance 2 runs the FP64 step 2-4% faster on the 8192^2 and
// 16384x8192 grids, and the split step 3-8% faster at tiles of 20 rows or
// more, but 4% slower at its best tile of 14 rows (launch_fused_t).

// Everything one wavefront carries while marching up its strip.
template <class P>
struct MarchState {
  typename P::Cell un, un2, ua, ub;  // rows r+1 (loaded), r+2 (distance 2), r-1, r-2
  Q4T<typename P::T> qm2, qm1, fymax_c, fx_c, fy_c;
  typename P::T wmax;
  // FP64: reciprocal of row r-1's density and whether quotients formed from it
  // can equal the IEEE ones (its density and reciprocal normal)
  double rm1;
  bool rokm1;
};

// One iteration over loaded row r.  STAGES is a compile-time set, so the
// steady-state iteration is one branch-free basic block in which the compiler
// can interleave the independent x-face and y-face Riemann chains.  COLS: the
// wavefront covers ghost columns (flips and ghost-column stores); ROWS: rows
// r - 2 .. r + 1 may be ghost rows or lie past the last row (row ghost map,
// flips and stores).  Without ROWS, rows r + PD <= ny + 1 and 2 <= r.  PF: 0
// loads nothing; with distance 1, PF = 1 loads row r + 1; with distance 2,
// PF = 1 moves the loaded row r + 1 up and PF = 2 also loads row r + 2.
template <int RS, class P, int STAGES, int PF, bool RATE, bool O32, bool COLS, bool ROWS, int PD>
__device__ __forceinline__ void march_row(const KParams &kp, const Grid &in, const Grid &out,
                                          MarchState<P> &m, int r, int c0, bool flipx,
                                          typename P::T dtdx, typename P::T dtdy,
                                          bool ghost_col, bool out_col, int col, int row_lo,
                                          int row_hi) {
  using T = typename P::T;
  using Cell = typename P::Cell;
  const SchemeT<T> &s = P::scheme(kp);
  const Cell uc = m.un;
  if constexpr (PD == 2 && PF >= 1) m.un = m.un2;
  if constexpr (PF == PD) {
    bool flipy = false;
    int rn = r + PD;
    if constexpr (ROWS)
      rn = ghost_map_uniform(min(r + PD, kp.ny + 3), kp.ny, kp.bcymin, kp.bcymax, flipy);
    Cell &dst = PD == 2 ? m.un2 : m.un;
    dst = P::template load_rc<O32>(in, rn, pin_offset((unsigned)c0 * P::kElem));
    if constexpr (COLS) {
      if (flipx) P::negate(dst, IU);
    }
    if constexpr (ROWS) {
      if (flipy) P::negate(dst, IV);
    }
  }
  Q4T<T> q0;
  [[maybe_unused]] double r0 = 0.0;
  [[maybe_unused]] bool rok0 = false, qok = true;
  if constexpr (P::kSharedRecip) {
    qok = P::prim_shared(kp, uc, q0, r0, rok0);
    if constexpr ((STAGES & kSlopes) == 0) {
      if (!wave_all(qok)) P::prim_ieee(kp, uc, q0);  // uniform
    }
  } else {
    P::prim(kp, uc, q0);
  }

  Q4T<T> fya = {}, fxa = {}, fymax = {};
  Cell unew;
  if constexpr ((STAGES & kSlopes) != 0) {
    const Q4T<T> qpx = shfl_dn1(m.qm1);
    const Q4T<T> qmx = shfl_up1(m.qm1);
    Q4T<T> dqX = {}, dqY = {};
    if constexpr ((RS & (kMinmodPow2 | kSlopeMinmod)) != 0) {  // slope_type == 1: no branch,
                                                                // no products by one
      slopes_xy<true>(s.slope_type, m.qm1, qpx, qmx, q0, m.qm2, dqX, dqY);
    } else if (s.slope_type != T(0.0)) {  // uniform: one scalar branch per row
      slopes_xy<false>(s.slope_type, m.qm1, qpx, qmx, q0, m.qm2, dqX, dqY);
    }
    SourcesT<T> src;
    if constexpr (P::kSharedRecip) {
      // one check for the quotients of row r's conversion and row a's source
      // terms; a wavefront with any lane outside the range redoes both, and
      // the y slopes that read row r, with the IEEE divisions
      src = trace_sources_shared(s, m.qm1, dqX, dqY, m.rm1, m.rokm1, qok);
      if (!wave_all(qok)) {  // uniform
        P::prim_ieee(kp, uc, q0);
        if constexpr ((RS & (kMinmodPow2 | kSlopeMinmod)) != 0)
          slopes_dir<true>(s.slope_type, m.qm1, q0, m.qm2, dqY);
        else if (s.slope_type != T(0.0))
          slopes_dir<false>(s.slope_type, m.qm1, q0, m.qm2, dqY);
        src = trace_sources(s, m.qm1, dqX, dqY);
      }
    } else {
      src = trace_sources(s, m.qm1, dqX, dqY);
    }
    Q4T<T> fymin;
    trace_face(s, m.qm1, dqY, src, dtdy, false, fymin);
    trace_face(s, m.qm1, dqY, src, dtdy, true, fymax);
    if constexpr ((STAGES & kYFace) != 0) {
      T f[4];
      riemann<RS>(s, m.fymax_c, fymin, IV, IU, f);
      for (int v = 0; v < 4; v++) fya.v[v] = f[v] * dtdy;
    }
    if constexpr ((STAGES & kXFace) != 0) {
      Q4T<T> fxmin, fxmax;
      trace_face(s, m.qm1, dqX, src, dtdx, false, fxmin);
      trace_face(s, m.qm1, dqX, src, dtdx, true, fxmax);
      const Q4T<T> ql = shfl_up1(fxmax);
      T f[4];
      riemann<RS>(s, ql, fxmin, IU, IV, f);
      for (int v = 0; v < 4; v++) fxa.v[v] = f[v] * dtdx;
    }
    if constexpr ((STAGES & kUpdate) != 0) {
      const Q4T<T> fxr = shfl_dn1(m.fx_c);
      unew = P::update(m.ub, m.fx_c, fxr, m.fy_c, fya);
    }
  }

  // stores last, so that the arithmetic above stays in one basic block
  // ghost cells of row r owned by this wave: materialise them in both arrays
  if constexpr (ROWS) {
    const bool row_ghost = r < 2 || (r > kp.ny + 1 && r <= kp.ny + 3);
    if ((ghost_col || row_ghost) && r >= row_lo && r <= row_hi) {
      const unsigned cb = pin_offset((unsigned)col * P::kElem);
      P::template store_rc<O32>(in, r, cb, uc);
      P::template store_rc<O32>(out, r, cb, uc);
    }
  } else if constexpr (COLS) {  // row r is an interior row: ghost columns only
    if (ghost_col && r >= row_lo && r <= row_hi) {
      const unsigned cb = pin_offset((unsigned)col * P::kElem);
      P::template store_rc<O32>(in, r, cb, uc);
      P::template store_rc<O32>(out, r, cb, uc);
    }
  }
  if constexpr ((STAGES & kUpdate) != 0) {
    const int b = r - 2;
    if (out_col && (!ROWS || b <= kp.ny + 1)) {
      P::template store_rc<O32>(out, b, pin_offset((unsigned)col * P::kElem), unew);
      if constexpr (RATE)  // next-step CFL rate of the new state
        m.wmax = Ops<T>::mx(m.wmax, cell_rate<P, (RS & kRiemannFloorSqrt) != 0,
                                              (RS & kMinmodPow2) != 0>(kp, unew));
    }
  }
  if constexpr ((STAGES & kYFace) != 0) {
    m.fx_c = fxa;
    m.fy_c = fya;
  }
  if constexpr ((STAGES & kSlopes) != 0) m.fymax_c = fymax;
  m.qm2 = m.qm1;
  m.qm1 = q0;
  if constexpr (P::kSharedRecip) {
    m.rm1 = r0;
    m.rokm1 = rok0;
  }
  m.ub = m.ua;
  m.ua = uc;
}

// One wavefront's march over strip `strip`, segment `seg` of one grid; returns
// the wave-local max of the next step's CFL rate when `want_rate`.
//
// Iteration k loads row r = j_base-2+k: k = 0,1 only convert; k = 2 adds the
// slopes of row j_base-1; k = 3 the first y and x faces; k = 4 .. seg_rows+2
// is the steady state (all stages, branch-free); k = seg_rows+3 finishes the
// last y face and update.
// Issue priority 3 - lvl (s_setprio takes an immediate).
__device__ __forceinline__ void set_prio_level(int lvl) {
  if (lvl <= 0)
    __builtin_amdgcn_s_setprio(3);
  else if (lvl == 1)
    __builtin_amdgcn_s_setprio(2);
  else if (lvl == 2)
    __builtin_amdgcn_s_setprio(1);
  else
    __builtin_amdgcn_s_setprio(0);
}

template <int RS, class P, bool RATE, bool O32, bool COLS, int PD>
__device__ __forceinline__ typename P::T march(const KParams &kp, const Grid &in, const Grid &out,
                                               double dt, int seg_rows, int nstrips, int strip,
                                               int seg, bool prio, bool go = true) {
  using T = typename P::T;
  const int lane = threadIdx.x & (kLanes - 1);

  const int nsegs = (kp.ny + seg_rows - 1) / seg_rows;
  const int i_base = 2 + kOutCols * strip;
  const int col = i_base - 2 + lane;
  const int j_base = 2 + seg_rows * seg;

  // A wavefront with COLS = false loads only interior columns and owns no
  // ghost column (interior_cols()).
  bool flipx = false;
  int c0 = col;
  bool col_owned = lane >= 2 && lane < 2 + kOutCols;
  bool ghost_col = false;
  if constexpr (COLS) {
    const int colc = min(col, kp.nx + 3);
    c0 = ghost_map(colc, kp.nx, kp.bcxmin, kp.bcxmax, flipx);
    const bool col_ghost = col < 2 || (col > kp.nx + 1 && col <= kp.nx + 3);
    col_owned = (lane >= 2 && lane < 2 + kOutCols && col <= kp.nx + 1) ||
                (strip == 0 && lane < 2) ||
                (strip == nstrips - 1 && col > kp.nx + 1 && col <= kp.nx + 3);
    ghost_col = col_ghost && col_owned;
  }
  const bool out_col = lane >= 2 && lane < 2 + kOutCols && (!COLS || col <= kp.nx + 1);
  // Ghost cells this wave writes: owned column, owned row, and a ghost column or
  // ghost row.  The owned rows of a segment are contiguous: [row_lo, row_hi].
  const int row_lo = seg == 0 ? 0 : j_base;
  const int row_hi = seg == nsegs - 1 ? kp.ny + 3 : min(j_base + seg_rows - 1, kp.ny + 1);
  const int wlo = col_owned ? row_lo : 1;  // empty range for columns not owned
  const int whi = col_owned ? row_hi : -1;

  MarchState<P> m;
  for (int d = 0; d < PD; d++) {  // rows j_base - 2 (and j_base - 1)
    bool flipy;
    const int rr =
        ghost_map_uniform(min(j_base - 2 + d, kp.ny + 3), kp.ny, kp.bcymin, kp.bcymax, flipy);
    typename P::Cell &dst = d == 0 ? m.un : m.un2;
    dst = P::template load_rc<O32>(in, rr, (unsigned)c0 * P::kElem);
    if constexpr (COLS) {
      if (flipx) P::negate(dst, IU);
    }
    if (flipy) P::negate(dst, IV);
  }
  // go = false (run loop ended): nothing is stored.  The check comes after the
  // first row's load is issued, so that its latency overlaps the loop-state read.
  if (!go) return T(0.0);
  const T dtdx = (T)(dt / kp.dx);
  const T dtdy = (T)(dt / kp.dy);
  m.ua = m.un;
  m.ub = m.un;
  m.qm2 = m.qm1 = m.fymax_c = m.fx_c = m.fy_c = Q4T<T>{};
  m.rm1 = 0.0;
  m.rokm1 = false;
  m.wmax = T(0.0);

  const int r0 = j_base - 2;
  // rows this segment updates: seg_rows, fewer in a top segment that ends at
  // row ny + 1 (its ghost rows ny + 2, ny + 3 are the last two loaded)
  const int rows = min(seg_rows, kp.ny + 2 - j_base);
#define E2D_ROW(ST, PF, RT, RW, k)                                                           \
  march_row<RS, P, ST, PF, RT, O32, COLS, RW, PD>(kp, in, out, m, r0 + (k), c0, flipx, dtdx,  \
                                                  dtdy, ghost_col, out_col, col, wlo, whi)
  E2D_ROW(0, PD, false, true, 0);
  E2D_ROW(0, PD, false, true, 1);
  E2D_ROW(kSlopes, PD, false, true, 2);
  E2D_ROW(kSlopes | kYFace | kXFace, PD, false, true, 3);
  // Steady state.  Iterations k <= ny + 1 - PD - r0 load row r + PD <= ny + 1,
  // and from k = 4 on row r >= 4: no row ghost handling.  They are unrolled
  // three times, so the rotating row buffers of MarchState are renamed
  // instead of copied.  Only the top segment has iterations past that bound.
  // The last load is of row r0 + rows + 3 (iteration rows + 3 - PD).
  const int k_rows = min(rows + 3 - PD, kp.ny + 1 - PD - r0);
  int k = 4;
  const int n1 = k_rows + 1;
  int k_next = prio ? k : INT_MAX;
  // prio: the issue priority falls from 3 to 0 over the steady rows, so that
  // of two wavefronts sharing a SIMD the one behind wins arbitration.  The
  // levels change at 1/2, 3/4 and 15/16 of the rows, checked before every row:
  // at equal priority the older wavefront wins, so the last level is kept
  // short.  Against quarters checked every third row this halves the gap
  // between the two wavefronts' ends (split 2.8 -> 1.7 us, FP64 6.3 -> 3.8 us
  // at 1024x1536, 14 rows) and takes 2% off the split step.
  const int t1 = (8 * n1 + 15) >> 4, t2 = (12 * n1 + 15) >> 4, t3 = (15 * n1 + 15) >> 4;
#define E2D_PRIO_AT(kk)                                               \
  if ((kk) >= k_next) {                                               \
    const int lvl = ((kk) >= t1) + ((kk) >= t2) + ((kk) >= t3);       \
    set_prio_level(lvl);                                              \
    k_next = lvl == 0 ? t1 : lvl == 1 ? t2 : lvl == 2 ? t3 : INT_MAX; \
  }
  for (; k + 2 <= k_rows; k += 3) {
    E2D_PRIO_AT(k);
    E2D_ROW(kAll, PD, RATE, false, k);
    E2D_PRIO_AT(k + 1);
    E2D_ROW(kAll, PD, RATE, false, k + 1);
    E2D_PRIO_AT(k + 2);
    E2D_ROW(kAll, PD, RATE, false, k + 2);
  }
  for (; k <= k_rows; ++k) {
    E2D_PRIO_AT(k);
    E2D_ROW(kAll, PD, RATE, false, k);
  }
#undef E2D_PRIO_AT
  for (; k <= rows + 3 - PD; ++k) E2D_ROW(kAll, PD, RATE, true, k);
  if constexpr (PD == 2) {  // row r0 + rows + 3 is loaded: move it up only
    for (; k <= rows + 2; ++k) E2D_ROW(kAll, 1, RATE, true, k);
  }
  E2D_ROW(kSlopes | kYFace | kUpdate, 0, RATE, true, rows + 3);
#undef E2D_ROW
  return m.wmax;
}

// Position of launch phase ph of a run loop (main.cpp:100,128-134): the time
// and step count before the step, the while condition, and dt = cfl / invDt
// (HydroRun.h:246) from the CFL rate of the step's input, clamped to tEnd.
// Every wavefront of the launch derives the same values from the same words.
struct LoopPos {
  double t, dt;
  int nstep;
  bool active;
};

__device__ __forceinline__ LoopPos loop_pos(const KParams &kp, const StepCtl *ctl, int ph) {
  LoopPos L;
  L.t = ctl->s[ph].t;
  L.nstep = ctl->s[ph].nstep;
  L.active = L.t < kp.tEnd && L.nstep < kp.nStepmax;
  // one vector load reads the 16 slots (16 lines) at once: lane k reads slot k.
  // The slots hold +0 or positive rates, so the max of the doubles is the max
  // of their bits.
  const int lane = threadIdx.x & (kLanes - 1);
  const double b =
      lane < kRateSlots ? __longlong_as_double((long long)ctl->rate[ph][lane * kSlotStride]) : 0.0;
  const long long mb = __double_as_longlong(wave_max(b));
  const unsigned mlo = __builtin_amdgcn_readfirstlane((unsigned)mb);
  const unsigned mhi = __builtin_amdgcn_readfirstlane((unsigned)(mb >> 32));
  const double invdt = __longlong_as_double((long long)(((unsigned long long)mhi << 32) | mlo));
  double dt = kp.cfl / invdt;
  if (L.t + dt > kp.tEnd) dt = kp.tEnd - L.t;
  L.dt = dt;
  return L;
}

// The step's one writer, the lanes of one wavefront: the state after the step
// (main.cpp:143, t += dt and nStep++; unchanged once the loop has ended) and
// the cleared rate entry of the launch after the next.  Nothing in this launch
// reads either.
__device__ __forceinline__ void loop_publish(StepCtl *ctl, int ph, const LoopPos &L, int lane) {
  const int p1 = ph == 2 ? 0 : ph + 1;
  const int p2 = ph == 0 ? 2 : ph - 1;
  if (lane < kRateSlots) ctl->rate[p2][lane * kSlotStride] = 0ull;
  if (lane == 0) {
    StepState n;
    n.t = L.active ? L.t + L.dt : L.t;
    n.nstep = L.active ? L.nstep + 1 : L.nstep;
    n.pad = 0;
    ctl->s[p1] = n;
    mirror_state(ctl->host_s, p1, n);
  }
}

// The rate slot the calling wavefront of launch phase ph folds into: slots are
// spread by global wavefront index.
__device__ __forceinline__ unsigned long long *rate_slot(StepCtl *ctl, int ph) {
  const int p1 = ph == 2 ? 0 : ph + 1;
  const unsigned wave = blockIdx.x * (kBlock / kLanes) + (threadIdx.x >> 6);
  return &ctl->rate[p1][(wave % kRateSlots) * kSlotStride];
}

// Folds the wavefront's maximum CFL rate into *target: one device-scope
// atomic per wavefront (the max is exact and order independent).  No barrier
// and no LDS merge: a wavefront that has finished its march leaves at once.
// A NaN rate never reaches a slot (m > 0 is false for it).
template <class T>
__device__ __forceinline__ void reduce_rate_wave(T wmax, unsigned long long *target) {
  const double m = (double)wave_max(wmax);
  if ((threadIdx.x & (kLanes - 1)) == 0 && target && m > 0.0) atomicMax(target, dbits(m));
}

// True when every column the wavefront of strip `strip` loads (60 strip ..
// 60 strip + 63) lies inside the domain: it then maps, flips and writes no
// ghost column.
__device__ __forceinline__ bool interior_cols(const KParams &kp, int strip) {
  const int c_lo = kOutCols * strip;
  return c_lo >= 2 && c_lo + kLanes - 1 <= kp.nx + 1;
}

// Wavefront -> (strip, segment, tile) of the fused step: the wavefronts of the
// interior strips first, then those of the edge strips (0 and nstrips - 1).
__device__ __forceinline__ void fused_wave(const FusedMap &f, int wave, int &strip, int &seg,
                                           int &tile) {
  if (wave < f.n_int) {
    const int ni = f.nstrips - 2;  // n_int > 0 only with more than two strips
    strip = 1 + wave % ni;
    seg = wave / ni;
    tile = f.tile;
  } else {
    const int w = wave - f.n_int;
    strip = w % f.nedge == 0 ? 0 : f.nstrips - 1;
    seg = w / f.nedge;
    tile = f.tile_e;
  }
}

// LOOP = false: explicit dt (HydroRun::godunov_unsplit); true: run-loop mode.
// O32: 32-bit row offsets (every plane at most 4 GiB).
template <int RS, class P, bool LOOP, bool O32, int PD>
__global__ __launch_bounds__(kBlock) void step_fused_kernel(KParams kp, Grid in, Grid out,
                                                            double dt_arg, StepCtl *ctl, int ph,
                                                            FusedMap fm) {
  using T = typename P::T;
  double dt = dt_arg;
  bool go = true;
  if constexpr (LOOP) {
    const LoopPos L = loop_pos(kp, ctl, ph);
    if (blockIdx.x == 0 && threadIdx.x < kLanes) loop_publish(ctl, ph, L, threadIdx.x);
    go = L.active;  // false once the loop has finished (uniform over the grid)
    dt = L.dt;
  }
  // the wave index is uniform: every row index and the row ghost map derived
  // from it are then scalar values and scalar control flow
  const int wave =
      __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kLanes) + (threadIdx.x >> 6));
  T wmax = T(0.0);
  if (wave < fm.nwaves) {
    int strip, seg, tile;
    fused_wave(fm, wave, strip, seg, tile);
    // in and out share their geometry (checked at launch): one pitch and one
    // plane stride let a row's offset serve both arrays
    Grid o = in;
    o.p = out.p;
    if (interior_cols(kp, strip))  // uniform
      wmax = march<RS, P, LOOP, O32, false, PD>(kp, in, o, dt, tile, fm.nstrips, strip, seg,
                                            fm.prio != 0, go);
    else
      wmax = march<RS, P, LOOP, O32, true, PD>(kp, in, o, dt, tile, fm.nstrips, strip, seg,
                                           fm.prio != 0, go);
  }
  // run loop: the CFL rate of the step's output, for the next launch's dt
  if constexpr (LOOP) {
    if (!go) return;
    reduce_rate_wave(wmax, rate_slot(ctl, ph));
  }
}

// ------------------------------------------------------- batched step kernel
//
// One launch advances every active instance of a batch by one step, each with
// its own dt.  Wavefront w works on the instance whose wave range contains w
// (binary search over wave_begin), on that instance's strip / segment; its
// in/out arrays follow the instance's own step parity.  Each instance has its
// own loop control rings (StepCtl), published by its first wavefront.
template <int RS, class P, bool O32>
__device__ __forceinline__ void batch_step_wave(const BatchInst *__restrict__ insts, int ninst,
                                                StepCtl *ctls, int ph, int nwaves,
                                                int seg_rows) {
  using T = typename P::T;
  const int wave =
      __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kLanes) + (threadIdx.x >> 6));
  T wmax = T(0.0);
  unsigned long long *target = nullptr;
  if (wave < nwaves) {
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {  // last instance with wave_begin <= wave
      const int mid = (lo + hi + 1) >> 1;
      if (insts[mid].wave_begin <= wave)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int inst = __builtin_amdgcn_readfirstlane(lo);  // wave-uniform: scalar loads of the table
    const BatchInst &bi = insts[inst];
    StepCtl *ctl = ctls + inst;
    // by-value copies: the march then reads its parameters from scalar
    // registers instead of reloading them around every store
    const KParams kp = bi.kp;
    const LoopPos L = loop_pos(kp, ctl, ph);
    const int w = wave - bi.wave_begin;
    if (w == 0) loop_publish(ctl, ph, L, threadIdx.x & (kLanes - 1));
    if (L.active) {
      const bool even = (L.nstep & 1) == 0;
      // U and U2 share their geometry (one allocation shape per instance)
      const Grid gin = even ? bi.U : bi.U2;
      Grid gout = gin;
      gout.p = even ? bi.U2.p : bi.U.p;
      const int strip = w % bi.nstrips, seg = w / bi.nstrips;
      if (interior_cols(kp, strip))
        wmax = march<RS, P, true, O32, false, 1>(kp, gin, gout, L.dt, seg_rows, bi.nstrips,
                                              strip, seg, false);
      else
        wmax = march<RS, P, true, O32, true, 1>(kp, gin, gout, L.dt, seg_rows, bi.nstrips,
                                             strip, seg, false);
      target = rate_slot(ctl, ph);
    }
  }
  reduce_rate_wave(wmax, target);
}

template <int RS, class P, bool O32>
__global__ __launch_bounds__(kBlock) void step_batch_kernel(const BatchInst *__restrict__ insts,
                                                            int ninst, StepCtl *ctls, int ph,
                                                            int nwaves, int seg_rows) {
  batch_step_wave<RS, P, O32>(insts, ninst, ctls, ph, nwaves, seg_rows);
}

// The split-precision step held to three wavefronts per SIMD (168 VGPRs, a few
// words spilled), where the unconstrained build (163-169 VGPRs with the
// 32-bit row offsets) may fit only two: 49.3 against 45.4 Gcell/s on the
// 512-instance sweep.
template <int RS, class P, bool O32>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
void step_batch_kernel_w3(const BatchInst *__restrict__ insts, int ninst, StepCtl *ctls, int ph,
                          int nwaves, int seg_rows) {
  batch_step_wave<RS, P, O32>(insts, ninst, ctls, ph, nwaves, seg_rows);
}

// ----------------------------------------------- persistent batched step kernel
//
// All the steps of a batch in one launch.  A wavefront reads the id of the XCD
// it runs on and takes work items from that XCD's queue (BatchQueue): step n
// of instance i, strip / segment w.  Before marching it waits until every
// item of step n - 1 of instance i has completed (per-instance count in
// `done`), so an instance advances as soon as its own previous step is done,
// with no grid-wide barrier between steps.  All items of an instance run on
// one XCD, whose L2 holds the rows its items write: a completed item's stores
// have reached that L2 (s_waitcnt before the count), and the next item drops
// its CU's L1 lines before reading.  The control rings are read and written
// with atomics, at the point where the rate atomics are performed.  Items are
// taken in step-major order, so an item only waits for items already taken
// by running wavefronts; every wait is bounded all the same (spin_limit polls
// of the count), and a wait that runs out sets counters->err and ends the launch.

__device__ __forceinline__ unsigned xcc_id() {
  unsigned x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return x & (kMaxXcds - 1);
}

__global__ void xcc_probe_kernel(unsigned *out) {
  const unsigned x = xcc_id();
  if (threadIdx.x == 0) out[blockIdx.x] = x;
}

// A word as the atomics left it: a device-scope atomic load, which the
// compiler issues as a vector load that bypasses the CU's caches (never a
// scalar load through the scalar cache, which atomics do not update).
__device__ __forceinline__ unsigned long long ctl_read(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned counter_read(unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ LoopPos loop_pos_coherent(const KParams &kp, StepCtl *ctl, int ph) {
  const int lane = threadIdx.x & (kLanes - 1);
  unsigned long long w0 = 0, w1 = 0;
  if (lane == 0) {
    unsigned long long *s = reinterpret_cast<unsigned long long *>(&ctl->s[ph]);
    w0 = ctl_read(s);
    w1 = ctl_read(s + 1);
  }
  const unsigned long long b =
      lane < kRateSlots ? ctl_read(&ctl->rate[ph][lane * kSlotStride]) : 0ull;
  LoopPos L;
  // readfirstlane returns an int: the halves are taken as unsigned, so that a
  // low half with its top bit set is not sign-extended into the high half
  const unsigned tlo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)w0);
  const unsigned thi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(w0 >> 32));
  L.t = __longlong_as_double((long long)(((unsigned long long)thi << 32) | tlo));
  L.nstep = (int)__builtin_amdgcn_readfirstlane((unsigned)w1);
  L.active = L.t < kp.tEnd && L.nstep < kp.nStepmax;
  const long long mb = __double_as_longlong(wave_max(__longlong_as_double((long long)b)));
  const unsigned mlo = __builtin_amdgcn_readfirstlane((unsigned)mb);
  const unsigned mhi = __builtin_amdgcn_readfirstlane((unsigned)(mb >> 32));
  const double invdt = __longlong_as_double((long long)(((unsigned long long)mhi << 32) | mlo));
  double dt = kp.cfl / invdt;
  if (L.t + dt > kp.tEnd) dt = kp.tEnd - L.t;
  L.dt = dt;
  return L;
}

__device__ __forceinline__ void loop_publish_coherent(StepCtl *ctl, int ph, const LoopPos &L,
                                                      int lane) {
  const int p1 = ph == 2 ? 0 : ph + 1;
  const int p2 = ph == 0 ? 2 : ph - 1;
  if (lane < kRateSlots) atomicExch(&ctl->rate[p2][lane * kSlotStride], 0ull);
  if (lane == 0) {
    const double t = L.active ? L.t + L.dt : L.t;
    const int n = L.active ? L.nstep + 1 : L.nstep;
    unsigned long long *s = reinterpret_cast<unsigned long long *>(&ctl->s[p1]);
    atomicExch(s, (unsigned long long)__double_as_longlong(t));
    atomicExch(s + 1, (unsigned long long)(unsigned)n);
  }
}

// The loop of one persistent wavefront.  Items are taken one at a time in
// queue order and marched at once, so the lowest item not yet completed is
// always held by a running wavefront whose waits are all on lower items.
template <int RS, class P, bool O32>
__device__ __forceinline__ void batch_persistent_loop(const BatchInst *__restrict__ insts,
                                                      StepCtl *ctls, int ph0, int steps,
                                                      const BatchQueue &q, BatchCounters *cnt,
                                                      unsigned *done, int seg_rows,
                                                      unsigned spin_limit) {
  using T = typename P::T;
  const int lane = threadIdx.x & (kLanes - 1);
  const int xq = q.xmap[xcc_id()];
  if (xq < 0) return;
  const unsigned per = (unsigned)q.per_step[xq];
  if (per == 0) return;
  const unsigned total = (unsigned)steps * per;
  const int b0 = q.xcd_begin[xq], b1 = q.xcd_begin[xq + 1];
  for (;;) {
    unsigned h = 0;
    if (lane == 0) h = atomicAdd(&cnt->head[xq][0], 1u);
    const unsigned it = __builtin_amdgcn_readfirstlane(h);
    if (it >= total) break;
    const int n = (int)(it / per);
    const int r = (int)(it - (unsigned)n * per);
    int lo = b0, hi = b1 - 1;
    while (lo < hi) {  // last position whose first item is at or before r
      const int mid = (lo + hi + 1) >> 1;
      if (q.wave_prefix[mid] <= r)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int pos = __builtin_amdgcn_readfirstlane(lo);
    const int inst = q.order[pos];
    const int w = r - q.wave_prefix[pos];
    const BatchInst &bi = insts[inst];
    const KParams kp = bi.kp;
    if (n > 0) {  // every item of step n - 1 of this instance has completed
      const unsigned target = (unsigned)n * (unsigned)bi.nstrips *
                              (unsigned)((kp.ny + seg_rows - 1) / seg_rows);
      bool ok = true;
      for (unsigned spins = 0;; ++spins) {
        unsigned d = 0, e = 0;
        if (lane == 0) {
          d = counter_read(&done[inst * kCounterStride]);
          e = counter_read(&cnt->err);
        }
        d = __builtin_amdgcn_readfirstlane(d);
        e = __builtin_amdgcn_readfirstlane(e);
        if (d >= target) break;
        if (e != 0u || spins >= spin_limit) {
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (!ok) {
        if (lane == 0) atomicOr(&cnt->err, 1u);
        break;
      }
    }
    // rows of instance i that this CU read in an earlier step may still sit in
    // its L1: drop them (the L2 of this XCD has the new ones).  The invalidate
    // takes agent scope (sc1): at workgroup scope (sc0) the CU's L1 counts as
    // coherent for its own workgroup and is left as it is.
    asm volatile("buffer_inv sc1" ::: "memory");
    const int ph = (ph0 + n) % 3;
    StepCtl *ctl = ctls + inst;
    const LoopPos L = loop_pos_coherent(kp, ctl, ph);
    if (w == 0) loop_publish_coherent(ctl, ph, L, lane);
    T wmax = T(0.0);
    if (L.active) {
      const bool even = (L.nstep & 1) == 0;
      const Grid gin = even ? bi.U : bi.U2;
      Grid gout = gin;
      gout.p = even ? bi.U2.p : bi.U.p;
      const int strip = w % bi.nstrips, seg = w / bi.nstrips;
      if (interior_cols(kp, strip))
        wmax = march<RS, P, true, O32, false, 1>(kp, gin, gout, L.dt, seg_rows, bi.nstrips,
                                                 strip, seg, false);
      else
        wmax = march<RS, P, true, O32, true, 1>(kp, gin, gout, L.dt, seg_rows, bi.nstrips,
                                                strip, seg, false);
      reduce_rate_wave(wmax, rate_slot(ctl, ph));
    }
    // this wavefront's stores and atomics are complete before it counts
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (lane == 0) atomicAdd(&done[inst * kCounterStride], 1u);
  }
}

template <int RS, class P, bool O32>
__global__ __launch_bounds__(kBlock) void step_batch_persistent_kernel(
    const BatchInst *__restrict__ insts, StepCtl *ctls, int ph0, int steps, BatchQueue q,
    BatchCounters *cnt, unsigned *done, int seg_rows, unsigned spin_limit) {
  batch_persistent_loop<RS, P, O32>(insts, ctls, ph0, steps, q, cnt, done, seg_rows, spin_limit);
}

// The split-precision loop held to three wavefronts per SIMD (168 VGPRs; the
// compiler spills a few words): a wavefront that waits on a queue or a
// dependency leaves two to issue, where the unconstrained build (175-180
// VGPRs) fits two per SIMD.  The FP64 loop would spill 13-23 words and stays
// at two.
template <int RS, class P, bool O32>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
void step_batch_persistent_kernel_w3(const BatchInst *__restrict__ insts, StepCtl *ctls, int ph0,
                                     int steps, BatchQueue q, BatchCounters *cnt,
                                     unsigned *done, int seg_rows, unsigned spin_limit) {
  batch_persistent_loop<RS, P, O32>(insts, ctls, ph0, steps, q, cnt, done, seg_rows, spin_limit);
}

// ----------------------------------------------------------- radial profile
// ComputeRadialProfileFunctor (ComputeRadialProfileFunctor.h:140-165): every
// cell of the array, ghosts included, adds its density to the bin of its
// distance from the box centre.  Counts are exact; the density sums are FP64
// atomics, so their last bits depend on the order, as in the reference.  Cells
// whose bin is >= nbins (ghost cells past the half-diagonal) are left out: the
// reference increments outside its histogram there.
constexpr int kProfileLdsBins = 1024;

template <class P>
__global__ void radial_profile_kernel(KParams kp, Grid U, RadialGeom geo,
                                      unsigned long long *counts, double *sums) {
  __shared__ double lsum[kProfileLdsBins];
  __shared__ unsigned int lcnt[kProfileLdsBins];
  const bool use_lds = geo.nbins <= kProfileLdsBins;
  if (use_lds)
    for (int b = threadIdx.x; b < geo.nbins; b += blockDim.x) {
      lsum[b] = 0.0;
      lcnt[b] = 0u;
    }
  __syncthreads();
  const long long n = (long long)kp.isize * kp.jsize;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(idx / kp.isize);
    const int i = (int)(idx - (long long)j * kp.isize);
    const double x = geo.xmin + kp.dx / 2 + (i - 2) * kp.dx;
    const double y = geo.ymin + kp.dy / 2 + (j - 2) * kp.dy;
    const double distance = sqrt((x - geo.cx) * (x - geo.cx) + (y - geo.cy) * (y - geo.cy));
    const int bin = (int)(distance / geo.max_radial_distance * geo.nbins);
    if (bin < 0 || bin >= geo.nbins) continue;
    const double rho = P::value(P::load(U, (long long)j * U.pitch + i), ID);
    if (use_lds) {
      atomicAdd(&lsum[bin], rho);
      atomicAdd(&lcnt[bin], 1u);
    } else {
      atomicAdd(&sums[bin], rho);
      atomicAdd(&counts[bin], 1ull);
    }
  }
  if (!use_lds) return;
  __syncthreads();
  for (int b = threadIdx.x; b < geo.nbins; b += blockDim.x)
    if (lcnt[b]) {
      atomicAdd(&sums[b], lsum[b]);
      atomicAdd(&counts[b], (unsigned long long)lcnt[b]);
    }
}

// -------------------------------------------------- conserved-variable sums
// Deterministic two-level FP64 sum: block b sums rows b, b+G, ... in a fixed tree.
constexpr int kSumBlocks = 512;

template <class P>
__global__ void sums_partial_kernel(KParams kp, Grid U, double *partials) {
  __shared__ double red[4][256];
  double acc[4] = {0, 0, 0, 0};
  for (int j = 2 + blockIdx.x; j <= kp.ny + 1; j += gridDim.x)
    for (int i = 2 + threadIdx.x; i <= kp.nx + 1; i += blockDim.x) {
      const typename P::Cell c = P::load(U, (long long)j * U.pitch + i);
      for (int v = 0; v < 4; v++) acc[v] += P::value(c, v);
    }
  for (int v = 0; v < 4; v++) red[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int v = 0; v < 4; v++) red[v][threadIdx.x] += red[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) partials[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void sums_final_kernel(const double *partials, int nblocks, double *out4) {
  __shared__ double red[4][256];
  double acc[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int v = 0; v < 4; v++) acc[v] += partials[b * 4 + v];
  for (int v = 0; v < 4; v++) red[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int v = 0; v < 4; v++) red[v][threadIdx.x] += red[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out4[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void fill_u32_kernel(unsigned *p, unsigned v) { *p = v; }

// Plain device copy (Kokkos::deep_copy of a whole state array, HydroRun.h:214),
// 16 bytes per lane per access.
__global__ void copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                            long long n2) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n2;
       k += (long long)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

// 32-bit row offsets reach every element of a plane of at most 4 GiB
bool off32_ok(const KParams &kp, const Grid &g) {
  return !kp.wide_addr && (unsigned long long)g.plane *
                                  (unsigned long long)grid_scalar_bytes(kp.prec) <=
                              (1ull << 32);
}

dim3 grid2d(int nx, int ny, dim3 blk) {
  return dim3((unsigned)((nx + blk.x - 1) / blk.x), (unsigned)((ny + blk.y - 1) / blk.y));
}

template <int RS, class P>
void launch_fused_t(const KParams &kp, Grid in, Grid out, double dt, StepCtl *ctl, int ph,
                    int seg_rows, hipStream_t st) {
  const FusedMap fm = fused_map(kp, seg_rows);
  const int waves_per_block = kBlock / kLanes;
  const int blocks = (fm.nwaves + waves_per_block - 1) / waves_per_block;
  // prefetch distance 2 where it measured faster: FP64 from 8-row tiles on,
  // the FP32 paths from 20-row tiles on
  const bool pd2 = seg_rows >= (std::is_same<P, P64>::value ? 8 : 20);
  if (ctl && off32_ok(kp, in) && off32_ok(kp, out)) {
    if (pd2)
      hipLaunchKernelGGL((step_fused_kernel<RS, P, true, true, 2>), dim3(blocks), dim3(kBlock), 0,
                         st, kp, in, out, dt, ctl, ph, fm);
    else
      hipLaunchKernelGGL((step_fused_kernel<RS, P, true, true, 1>), dim3(blocks), dim3(kBlock), 0,
                         st, kp, in, out, dt, ctl, ph, fm);
  } else if constexpr ((RS & kMinmodPow2) == 0) {  // the minmod kinds have the branch above only
    if (ctl)
      hipLaunchKernelGGL((step_fused_kernel<RS, P, true, false, 1>), dim3(blocks), dim3(kBlock),
                         0, st, kp, in, out, dt, ctl, ph, fm);
    else
      hipLaunchKernelGGL((step_fused_kernel<RS, P, false, false, 1>), dim3(blocks), dim3(kBlock),
                         0, st, kp, in, out, dt, ctl, ph, fm);
  }
}

template <class P>
void launch_fused_p(const KParams &kp, Grid in, Grid out, double dt, StepCtl *ctl, int ph,
                    int seg_rows, hipStream_t st) {
  if constexpr (std::is_same<P, P64>::value) {
    if (kp.s.floor_sqrt && kp.s.riemann != 0) {  // HLL / HLLC with the short square root
      // minmod slopes and power-of-two cell sizes built in: run-loop launches
      // with 32-bit row offsets only
      if (kp.s.slope_type == 1.0 && kp.rd.pow2x && kp.rd.pow2y && ctl && off32_ok(kp, in) &&
          off32_ok(kp, out)) {
        if (kp.s.riemann == 2)
          launch_fused_t<2 | kRiemannFloorSqrt | kMinmodPow2, P>(kp, in, out, dt, ctl, ph,
                                                                 seg_rows, st);
        else
          launch_fused_t<1 | kRiemannFloorSqrt | kMinmodPow2, P>(kp, in, out, dt, ctl, ph,
                                                                 seg_rows, st);
        return;
      }
      if (kp.s.riemann == 2)
        launch_fused_t<2 | kRiemannFloorSqrt, P>(kp, in, out, dt, ctl, ph, seg_rows, st);
      else
        launch_fused_t<1 | kRiemannFloorSqrt, P>(kp, in, out, dt, ctl, ph, seg_rows, st);
      return;
    }
  }
  if (kp.s.riemann == 2)
    launch_fused_t<2, P>(kp, in, out, dt, ctl, ph, seg_rows, st);
  else if (kp.s.riemann == 1)
    launch_fused_t<1, P>(kp, in, out, dt, ctl, ph, seg_rows, st);
  else
    launch_fused_t<0, P>(kp, in, out, dt, ctl, ph, seg_rows, st);
}

}  // namespace

int grid_planes(int) { return 4; }
int grid_scalar_bytes(int precision) { return precision == E2D_PRECISION_FP32 ? 4 : 8; }

KParams make_kparams(const e2d_params &p, int precision) {
  KParams k{};
  k.s.gamma0 = p.gamma0;
  k.s.gamma6 = p.gamma6;
  k.s.smallr = p.smallr;
  k.s.smallc = p.smallc;
  k.s.smallp = p.smallp;
  k.s.smallpp = p.smallpp;
  k.s.slope_type = p.slope_type;
  k.s.entho = 1.0 / (p.gamma0 - 1.0);
  k.s.smallc2 = p.smallc * p.smallc;
  // the short square root for the wave speeds (arguments >= smallc^2) and
  // the CFL rate (arguments >= about smallc^2 while d smallp is normal)
  k.s.floor_sqrt =
      k.s.smallc2 >= std::ldexp(1.0, -759) && p.smallr * p.smallp >= std::ldexp(1.0, -1000)
          ? 1
          : 0;
  k.s.riemann = p.honor_riemann ? p.riemannSolverType : 2;
  k.sf.gamma0 = (float)k.s.gamma0;
  k.sf.gamma6 = (float)k.s.gamma6;
  k.sf.smallr = (float)k.s.smallr;
  k.sf.smallc = (float)k.s.smallc;
  k.sf.smallp = (float)k.s.smallp;
  k.sf.smallpp = (float)k.s.smallpp;
  k.sf.slope_type = (float)k.s.slope_type;
  k.sf.entho = (float)k.s.entho;
  k.sf.smallc2 = (float)k.s.smallc2;
  k.sf.floor_sqrt = 0;
  k.sf.riemann = k.s.riemann;
  k.prec = precision;
  k.nx = p.nx;
  k.ny = p.ny;
  k.isize = p.isize;
  k.jsize = p.jsize;
  k.bcxmin = p.boundary_type_xmin;
  k.bcxmax = p.boundary_type_xmax;
  k.bcymin = p.boundary_type_ymin;
  k.bcymax = p.boundary_type_ymax;
  k.nStepmax = p.nStepmax;
  k.dx = p.dx;
  k.dy = p.dy;
  k.dxf = (float)p.dx;
  k.dyf = (float)p.dy;
  k.cfl = p.cfl;
  k.tEnd = p.tEnd;
  k.kin_frac = (float)p.kin_frac;
  // exact powers of two (normal, with a normal reciprocal) in both FP64 and FP32
  auto pow2 = [](double h) {
    int e = 0;
    return h > 0.0 && std::frexp(h, &e) == 0.5 && e > -125 && e < 126;
  };
  k.rd = RateT<double>{p.dx, p.dy, 1.0 / p.dx, 1.0 / p.dy, pow2(p.dx), pow2(p.dy)};
  k.rf = RateT<float>{k.dxf, k.dyf, (float)(1.0 / p.dx), (float)(1.0 / p.dy), pow2(p.dx),
                      pow2(p.dy)};
  return k;
}

#define E2D_DISPATCH_P(kp, KERNEL, ...)                   \
  do {                                                    \
    if ((kp).prec == E2D_PRECISION_FP32_SPLIT)            \
      hipLaunchKernelGGL(KERNEL<PSplit>, __VA_ARGS__);    \
    else if ((kp).prec == E2D_PRECISION_FP32)             \
      hipLaunchKernelGGL(KERNEL<P32>, __VA_ARGS__);       \
    else                                                  \
      hipLaunchKernelGGL(KERNEL<P64>, __VA_ARGS__);       \
  } while (0)

hipError_t launch_init(const KParams &kp, const e2d_params &p, Grid U, unsigned int *count_dev,
                       hipStream_t st) {
  const dim3 blk(64, 4);
  E2D_DISPATCH_P(kp, init_kernel, grid2d(p.isize, p.jsize, blk), blk, 0, st, p, U, count_dev);
  return hipGetLastError();
}

hipError_t launch_blast_energy(const KParams &kp, const e2d_params &p, Grid U, double e_inside,
                               hipStream_t st) {
  const dim3 blk(64, 4);
  E2D_DISPATCH_P(kp, blast_energy_kernel, grid2d(p.isize, p.jsize, blk), blk, 0, st, p, U,
                 e_inside);
  return hipGetLastError();
}

hipError_t launch_boundaries(const KParams &kp, Grid U, hipStream_t st) {
  const int n = 4 * kp.isize + 4 * kp.ny;
  E2D_DISPATCH_P(kp, boundaries_kernel, dim3((n + 255) / 256), dim3(256), 0, st, kp, U);
  return hipGetLastError();
}

hipError_t launch_dt_reduce(const KParams &kp, Grid U, unsigned long long *invdt_bits,
                            hipStream_t st) {
  const int blocks = kp.ny < 2048 ? kp.ny : 2048;
  E2D_DISPATCH_P(kp, dt_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, st, kp, U,
                 invdt_bits);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned *p, unsigned v, hipStream_t st) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(1), 0, st, p, v);
  return hipGetLastError();
}

hipError_t launch_loop_begin(const KParams &kp, Grid in, Grid other, int bc, StepCtl *ctl,
                             double t, int nstep, double *partials, hipStream_t st) {
  E2D_DISPATCH_P(kp, loop_begin_kernel, dim3(kBeginBlocks), dim3(kBeginThreads), 0, st, kp, in,
                 other, bc, ctl, t, nstep, partials, &ctl->arrivals);
  return hipGetLastError();
}

hipError_t launch_kinetic_count(const KParams &kp, Grid U, unsigned long long *count,
                                hipStream_t st) {
  const int blocks = kp.ny < 2048 ? kp.ny : 2048;
  E2D_DISPATCH_P(kp, kinetic_count_kernel, dim3((unsigned)blocks), dim3(256), 0, st, kp, U,
                 count);
  return hipGetLastError();
}

hipError_t launch_prim(const KParams &kp, Grid U, Grid Q, hipStream_t st) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(prim_kernel, grid2d(kp.isize, kp.jsize, blk), blk, 0, st, kp, U, Q);
  return hipGetLastError();
}

hipError_t launch_fluxes(const KParams &kp, Grid Q, Grid Fx, Grid Fy, double dtdx, double dtdy,
                         hipStream_t st) {
  const dim3 blk(64, 4);
  const dim3 grd = grid2d(kp.nx + 1, kp.ny + 1, blk);
  if (kp.s.riemann == 2)
    hipLaunchKernelGGL(fluxes_kernel<2>, grd, blk, 0, st, kp, Q, Fx, Fy, dtdx, dtdy);
  else if (kp.s.riemann == 1)
    hipLaunchKernelGGL(fluxes_kernel<1>, grd, blk, 0, st, kp, Q, Fx, Fy, dtdx, dtdy);
  else
    hipLaunchKernelGGL(fluxes_kernel<0>, grd, blk, 0, st, kp, Q, Fx, Fy, dtdx, dtdy);
  return hipGetLastError();
}

hipError_t launch_update(const KParams &kp, Grid U, Grid Fx, Grid Fy, hipStream_t st) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(update_kernel, grid2d(kp.nx, kp.ny, blk), blk, 0, st, kp, U, Fx, Fy);
  return hipGetLastError();
}

hipError_t launch_slopes(const KParams &kp, Grid Q, Grid Sx, Grid Sy, hipStream_t st) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(slopes_kernel, grid2d(kp.isize - 2, kp.jsize - 2, blk), blk, 0, st, kp, Q, Sx,
                     Sy);
  return hipGetLastError();
}

hipError_t launch_trace_fluxes(const KParams &kp, int dir, Grid Q, Grid Sx, Grid Sy, Grid F,
                               double dtdx, double dtdy, hipStream_t st) {
  const dim3 blk(64, 4);
  const dim3 grd = grid2d(kp.nx + 1, kp.ny + 1, blk);
#define E2D_TF(RS, D) \
  hipLaunchKernelGGL((trace_fluxes_kernel<RS, D>), grd, blk, 0, st, kp, Q, Sx, Sy, F, dtdx, dtdy)
  if (dir == 0) {
    if (kp.s.riemann == 2) E2D_TF(2, 0);
    else if (kp.s.riemann == 1) E2D_TF(1, 0);
    else E2D_TF(0, 0);
  } else {
    if (kp.s.riemann == 2) E2D_TF(2, 1);
    else if (kp.s.riemann == 1) E2D_TF(1, 1);
    else E2D_TF(0, 1);
  }
#undef E2D_TF
  return hipGetLastError();
}

hipError_t launch_update_dir(const KParams &kp, int dir, Grid U, Grid F, hipStream_t st) {
  const dim3 blk(64, 4);
  if (dir == 0)
    hipLaunchKernelGGL(update_dir_kernel<0>, grid2d(kp.nx, kp.ny, blk), blk, 0, st, kp, U, F);
  else
    hipLaunchKernelGGL(update_dir_kernel<1>, grid2d(kp.nx, kp.ny, blk), blk, 0, st, kp, U, F);
  return hipGetLastError();
}

hipError_t launch_fluxes_atomic(const KParams &kp, Grid Q, Grid U, double dtdx, double dtdy,
                                hipStream_t st) {
  const dim3 blk(64, 4);
  const dim3 grd = grid2d(kp.nx + 1, kp.ny + 1, blk);
  if (kp.s.riemann == 2)
    hipLaunchKernelGGL(fluxes_atomic_kernel<2>, grd, blk, 0, st, kp, Q, U, dtdx, dtdy);
  else if (kp.s.riemann == 1)
    hipLaunchKernelGGL(fluxes_atomic_kernel<1>, grd, blk, 0, st, kp, Q, U, dtdx, dtdy);
  else
    hipLaunchKernelGGL(fluxes_atomic_kernel<0>, grd, blk, 0, st, kp, Q, U, dtdx, dtdy);
  return hipGetLastError();
}

hipError_t launch_step_fused(const KParams &kp, Grid in, Grid out, double dt, StepCtl *ctl,
                             int ph, int seg_rows, hipStream_t st) {
  if (ctl && (ph < 0 || ph > 2)) return hipErrorInvalidValue;
  if (in.pitch != out.pitch || in.plane != out.plane) return hipErrorInvalidValue;
  if (seg_rows < 1) seg_rows = 1;
  if (kp.prec == E2D_PRECISION_FP32_SPLIT)
    launch_fused_p<PSplit>(kp, in, out, dt, ctl, ph, seg_rows, st);
  else if (kp.prec == E2D_PRECISION_FP32)
    launch_fused_p<P32>(kp, in, out, dt, ctl, ph, seg_rows, st);
  else
    launch_fused_p<P64>(kp, in, out, dt, ctl, ph, seg_rows, st);
  return hipGetLastError();
}

FusedMap fused_map(const KParams &kp, int seg_rows) {
  FusedMap f{};
  f.tile = seg_rows;
  f.tile_e = std::max(1, (seg_rows + 4) * 10 / 11 - 4);
  f.nstrips = (kp.nx + kOutCols - 1) / kOutCols;
  f.nedge = std::min(f.nstrips, 2);
  const int nsegs_i = (kp.ny + f.tile - 1) / f.tile;
  const int nsegs_e = (kp.ny + f.tile_e - 1) / f.tile_e;
  f.n_int = (f.nstrips - f.nedge) * nsegs_i;
  f.nwaves = f.n_int + f.nedge * nsegs_e;
  f.prio = kp.simds > 0 && f.nwaves > kp.simds && f.nwaves <= 2 * kp.simds ? 1 : 0;
  return f;
}

int batch_waves(const KParams &kp, int seg_rows, int *nstrips) {
  const int ns = (kp.nx + kOutCols - 1) / kOutCols;
  if (nstrips) *nstrips = ns;
  return ns * ((kp.ny + seg_rows - 1) / seg_rows);
}

hipError_t launch_xcc_probe(unsigned *out, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(xcc_probe_kernel, dim3((unsigned)blocks), dim3(kLanes), 0, st, out);
  return hipGetLastError();
}

// One block of four wavefronts per slot the kernel's registers leave on a CU,
// times the CUs: every block is resident from the start, and none waits for
// another to exit.
template <class K>
hipError_t launch_persistent(K kernel, int ncu, hipStream_t st, const BatchInst *insts_dev,
                             StepCtl *ctls_dev, int ph0, int steps, const BatchQueue &q,
                             BatchCounters *counters, unsigned *done, int seg_rows,
                             unsigned spin_limit) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  hipLaunchKernelGGL(kernel, dim3(per_cu * ncu), dim3(kBlock), 0, st, insts_dev, ctls_dev, ph0,
                     steps, q, counters, done, seg_rows, spin_limit);
  return hipGetLastError();
}

hipError_t launch_step_batch_persistent(const KParams &kp0, const BatchInst *insts_dev, int ninst,
                                        StepCtl *ctls_dev, int ph0, int steps, BatchQueue q,
                                        BatchCounters *counters, unsigned *done, int seg_rows,
                                        unsigned spin_limit, int ncu, hipStream_t st,
                                        bool minmod) {
  (void)ninst;
  if (ph0 < 0 || ph0 > 2 || steps <= 0 || ncu <= 0) return hipErrorInvalidValue;
#define E2D_PBATCH_ARGS ncu, st, insts_dev, ctls_dev, ph0, steps, q, counters, done, seg_rows, \
                        spin_limit
#define E2D_PBATCH_LAUNCH(K, RS, P)                                            \
  return kp0.wide_addr ? launch_persistent(K<RS, P, false>, E2D_PBATCH_ARGS) \
                       : launch_persistent(K<RS, P, true>, E2D_PBATCH_ARGS)
  if (kp0.prec == E2D_PRECISION_FP32_SPLIT) {
    if (kp0.s.riemann == 2) E2D_PBATCH_LAUNCH(step_batch_persistent_kernel_w3, 2, PSplit);
    if (kp0.s.riemann == 1) E2D_PBATCH_LAUNCH(step_batch_persistent_kernel_w3, 1, PSplit);
    E2D_PBATCH_LAUNCH(step_batch_persistent_kernel_w3, 0, PSplit);
  }
  if (kp0.prec == E2D_PRECISION_FP32) {
    if (kp0.s.riemann == 2) E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 2, P32);
    if (kp0.s.riemann == 1) E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 1, P32);
    E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 0, P32);
  }
  if (kp0.s.floor_sqrt && kp0.s.riemann != 0) {
    if (minmod && kp0.s.riemann == 2)
      E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 2 | kRiemannFloorSqrt | kSlopeMinmod, P64);
    if (minmod)
      E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 1 | kRiemannFloorSqrt | kSlopeMinmod, P64);
    if (kp0.s.riemann == 2)
      E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 2 | kRiemannFloorSqrt, P64);
    E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 1 | kRiemannFloorSqrt, P64);
  }
  if (kp0.s.riemann == 2) E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 2, P64);
  if (kp0.s.riemann == 1) E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 1, P64);
  E2D_PBATCH_LAUNCH(step_batch_persistent_kernel, 0, P64);
#undef E2D_PBATCH_LAUNCH
#undef E2D_PBATCH_ARGS
}

hipError_t launch_step_batch(const KParams &kp0, const BatchInst *insts_dev, int ninst,
                             StepCtl *ctls_dev, int ph, int nwaves, int seg_rows, hipStream_t st,
                             bool minmod) {
  if (ph < 0 || ph > 2) return hipErrorInvalidValue;
  const int blocks = (nwaves + kBlock / kLanes - 1) / (kBlock / kLanes);
#define E2D_BATCH_LAUNCH(K, RS, P)                                                            \
  do {                                                                                        \
    if (kp0.wide_addr)                                                                        \
      hipLaunchKernelGGL((K<RS, P, false>), dim3(blocks), dim3(kBlock), 0, st, insts_dev,     \
                         ninst, ctls_dev, ph, nwaves, seg_rows);                               \
    else                                                                                      \
      hipLaunchKernelGGL((K<RS, P, true>), dim3(blocks), dim3(kBlock), 0, st, insts_dev,      \
                         ninst, ctls_dev, ph, nwaves, seg_rows);                               \
  } while (0)
  if (kp0.prec == E2D_PRECISION_FP32_SPLIT) {
    if (kp0.s.riemann == 2) E2D_BATCH_LAUNCH(step_batch_kernel_w3, 2, PSplit);
    else if (kp0.s.riemann == 1) E2D_BATCH_LAUNCH(step_batch_kernel_w3, 1, PSplit);
    else E2D_BATCH_LAUNCH(step_batch_kernel_w3, 0, PSplit);
  } else if (kp0.prec == E2D_PRECISION_FP32) {
    if (kp0.s.riemann == 2) E2D_BATCH_LAUNCH(step_batch_kernel, 2, P32);
    else if (kp0.s.riemann == 1) E2D_BATCH_LAUNCH(step_batch_kernel, 1, P32);
    else E2D_BATCH_LAUNCH(step_batch_kernel, 0, P32);
  } else if (kp0.s.floor_sqrt && kp0.s.riemann != 0 && minmod) {
    if (kp0.s.riemann == 2)
      E2D_BATCH_LAUNCH(step_batch_kernel, 2 | kRiemannFloorSqrt | kSlopeMinmod, P64);
    else
      E2D_BATCH_LAUNCH(step_batch_kernel, 1 | kRiemannFloorSqrt | kSlopeMinmod, P64);
  } else if (kp0.s.floor_sqrt && kp0.s.riemann != 0) {
    if (kp0.s.riemann == 2) E2D_BATCH_LAUNCH(step_batch_kernel, 2 | kRiemannFloorSqrt, P64);
    else E2D_BATCH_LAUNCH(step_batch_kernel, 1 | kRiemannFloorSqrt, P64);
  } else {
    if (kp0.s.riemann == 2) E2D_BATCH_LAUNCH(step_batch_kernel, 2, P64);
    else if (kp0.s.riemann == 1) E2D_BATCH_LAUNCH(step_batch_kernel, 1, P64);
    else E2D_BATCH_LAUNCH(step_batch_kernel, 0, P64);
  }
#undef E2D_BATCH_LAUNCH
  return hipGetLastError();
}

int sums_partials_needed(const KParams &) { return kSumBlocks * 4; }

hipError_t launch_sums(const KParams &kp, Grid U, double *partials, double *out4, hipStream_t st) {
  E2D_DISPATCH_P(kp, sums_partial_kernel, dim3(kSumBlocks), dim3(256), 0, st, kp, U, partials);
  hipLaunchKernelGGL(sums_final_kernel, dim3(1), dim3(256), 0, st, partials, kSumBlocks, out4);
  return hipGetLastError();
}

hipError_t launch_radial_profile(const KParams &kp, Grid U, const RadialGeom &geo,
                                 unsigned long long *counts, double *sums, hipStream_t st) {
  if (geo.nbins < 1) return hipErrorInvalidValue;
  hipError_t e = hipMemsetAsync(counts, 0, sizeof(unsigned long long) * geo.nbins, st);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(sums, 0, sizeof(double) * geo.nbins, st);
  if (e != hipSuccess) return e;
  const long long n = (long long)kp.isize * kp.jsize;
  long long blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  E2D_DISPATCH_P(kp, radial_profile_kernel, dim3((unsigned)blocks), dim3(256), 0, st, kp, U, geo,
                 counts, sums);
  return hipGetLastError();
}

hipError_t launch_copy(const KParams &kp, Grid src, Grid dst, hipStream_t st) {
  // every plane starts on a 512-byte boundary, so the array is whole 16-byte units
  const long long bytes = src.plane * grid_planes(kp.prec) * grid_scalar_bytes(kp.prec);
  const long long n2 = bytes / 16;
  long long blocks = (n2 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                     static_cast<const double2 *>(src.p), static_cast<double2 *>(dst.p), n2);
  return hipGetLastError();
}

}  // namespace e2d
