// numerics.hpp -- per-cell / per-face arithmetic of the MUSCL-Hancock scheme.
//
// Templated on the arithmetic type T.  With T = double each function evaluates
// the same IEEE operations, in the same order, as the reference's inline
// functions (pkestene/euler2d_kokkos src/HydroBaseFunctor.h), so that the FP64
// path reproduces the reference bit for bit: this file must be compiled without
// FMA contraction (-ffp-contract=off and the pragma below) and without
// fast-math.  Restructurings that keep every FP64 result bit-identical are
// marked where they happen (shared source terms, one star state per HLLC).
// With T = float (the split-precision path) division and square root use the
// hardware reciprocal / square-root instructions (about 1 ulp), which is
// where that path saves most of its instructions.
#pragma once

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace e2d {

enum { ID = 0, IP = 1, IU = 2, IV = 3 };

// Division, square root, min/max per arithmetic type.
template <class T>
struct Ops;

// mad(a, b, c) = a * b + c: for FP64 the product and the sum are rounded
// separately, exactly as the reference writes the expression (the FP64 path
// must keep every rounding); for FP32 it is one fused multiply-add.  Callers
// pass the operands so that the FP64 form is the reference's own expression
// tree (x - a * b is written mad(-a, b, x): negating a factor is exact).
template <>
struct Ops<double> {
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double mad(double a, double b, double c) { return a * b + c; }
  static __device__ __forceinline__ double sqroot(double x) { return sqrt(x); }
  static __device__ __forceinline__ double mx(double a, double b) { return fmax(a, b); }
  static __device__ __forceinline__ double mn(double a, double b) { return fmin(a, b); }
  static __device__ __forceinline__ double ab(double a) { return fabs(a); }
};

// Square root of x >= smallc^2 for the wave-speed estimates.  The compiler's
// correctly rounded FP64 sqrt first scales x < 2^-767 by 2^256, then runs
// rsq + one Goldschmidt step + two Newton corrections, unscales, and returns
// x itself for +-0 and +inf.  For x in [2^-767, DBL_MAX] the scaling and the
// special cases are identities, so the same core sequence, operation for
// operation, gives the same bits; the host enables it only when smallc^2 (the
// floor of every argument) is >= 2^-759.  A state with infinite pressure gets
// NaN here where sqrt returns inf.
__device__ __forceinline__ double sqrt_core(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  double d = __builtin_fma(-g, g, x);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return g;
}

template <>
struct Ops<float> {
  static __device__ __forceinline__ float div(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float mad(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
  }
  static __device__ __forceinline__ float sqroot(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float mx(float a, float b) { return fmaxf(a, b); }
  static __device__ __forceinline__ float mn(float a, float b) { return fminf(a, b); }
  static __device__ __forceinline__ float ab(float a) { return fabsf(a); }
};

// Scheme constants, precomputed on the host in double (entho = 1/(gamma0-1) is
// the same IEEE division the reference performs inside every Riemann call)
// and rounded once for the float path.
template <class T>
struct SchemeT {
  T gamma0, gamma6, smallr, smallc, smallp, smallpp, slope_type, entho;
  T smallc2;    // smallc * smallc
  int riemann;  // 0 approx, 1 hll, 2 hllc (already resolved from honor_riemann)
  int floor_sqrt;  // smallc2 >= 2^-759 and smallr*smallp >= 2^-1000: sound speeds take the short sqrt
};
using Scheme = SchemeT<double>;

// sqrt(x) for x >= s.smallc2 (the wave-speed estimates of the Riemann
// solvers); FS selects the short sequence (FP64 only, when s.floor_sqrt)
template <bool FS, class T>
__device__ __forceinline__ T floored_sqrt(T x) {
  if constexpr (FS)
    return sqrt_core(x);
  else
    return Ops<T>::sqroot(x);
}

template <class T>
struct Q4T {
  T v[4];
};
using Q4 = Q4T<double>;

// ------------------------------------------------ FP64 quotients sharing d
// The compiler's IEEE FP64 division a / d is div_scale(d), rcp, two Newton
// steps, m = a r, e = a - d m, q = m + e r (div_fmas), div_fixup.  For the
// quotients of one cell by its density the reciprocal is formed once, by the
// same operations, and each quotient takes the three steps after it.  The
// scaling and fix-up steps act only where d or 1/d is not normal, where the
// residual e is denormal (a below about 2^-969), or where q is not normal: the
// class tests below catch each of these, and a wavefront with any such lane
// takes the IEEE division for all of its lanes.  A zero quotient takes the
// sign of m, as div_fixup gives it.  Same bits; 27 instead of 33 VALU
// instructions for the three quotients of prim.  Each fallback is a branch
// that splits the march's steady row, and a third branch per row measured
// slower than no sharing at all; so quotients by one density that a row forms
// later (the trace's source terms, the CFL rate's gamma p / d) join the check
// of that density's conversion instead of adding one.

__device__ __forceinline__ double recip_newton(double d) {
  const double r0 = __builtin_amdgcn_rcp(d);
  const double r1 = __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
  return __builtin_fma(r1, __builtin_fma(-d, r1, 1.0), r1);
}

constexpr int kClassNormal = (1 << 3) | (1 << 8);
constexpr int kClassZeroOrNormal = kClassNormal | (1 << 5) | (1 << 6);

__device__ __forceinline__ double quot_shared(double a, double d, double r, bool &ok) {
  const double m = a * r;
  const double e = __builtin_fma(-d, m, a);
  const double q = __builtin_fma(e, r, m);
  ok = ok && __builtin_amdgcn_class(q, kClassZeroOrNormal) &&
       __builtin_amdgcn_class(e, kClassZeroOrNormal);
  return __builtin_copysign(q, m);
}

// true in every lane when no active lane of the wavefront has !ok
__device__ __forceinline__ bool wave_all(bool ok) {
  return __builtin_amdgcn_ballot_w64(!ok) == 0;
}

// the rest of computePrimitives once the three quotients are formed
__device__ __forceinline__ void prim_finish(const SchemeT<double> &s, double d, double ux,
                                            double uy, double e1, Q4T<double> &q) {
  const double eken = 0.5 * (uy * uy + ux * ux);
  const double e = e1 - eken;
  const double p = fmax((s.gamma0 - 1.0) * d * e, d * s.smallp);
  q.v[ID] = d;
  q.v[IP] = p;
  q.v[IU] = ux;
  q.v[IV] = uy;
}

// computePrimitives for FP64 with the three quotients formed from one
// reciprocal and no fallback.  Returns whether this lane's quotients are the
// IEEE ones; r is the reciprocal and rok whether other quotients by d formed
// from it can be (d and r normal).
__device__ __forceinline__ bool prim_fp64_shared(const SchemeT<double> &s, double u0, double u1,
                                                 double u2, double u3, Q4T<double> &q, double &r,
                                                 bool &rok) {
  const double d = fmax(u0, s.smallr);
  r = recip_newton(d);
  rok = __builtin_amdgcn_class(d, kClassNormal) && __builtin_amdgcn_class(r, kClassNormal);
  bool ok = rok;
  const double ux = quot_shared(u2, d, r, ok);
  const double uy = quot_shared(u3, d, r, ok);
  const double e1 = quot_shared(u1, d, r, ok);
  prim_finish(s, d, ux, uy, e1, q);
  return ok;
}

// computePrimitives for FP64: the shared reciprocal, and the IEEE divisions
// for a wavefront with any lane outside its range
__device__ __forceinline__ void prim_fp64(const SchemeT<double> &s, double u0, double u1,
                                          double u2, double u3, Q4T<double> &q) {
  const double d = fmax(u0, s.smallr);
  const double r = recip_newton(d);
  bool ok = __builtin_amdgcn_class(d, kClassNormal) && __builtin_amdgcn_class(r, kClassNormal);
  double ux = quot_shared(u2, d, r, ok);
  double uy = quot_shared(u3, d, r, ok);
  double e1 = quot_shared(u1, d, r, ok);
  if (!wave_all(ok)) {  // uniform
    ux = u2 / d;
    uy = u3 / d;
    e1 = u1 / d;
  }
  prim_finish(s, d, ux, uy, e1, q);
}

// computePrimitives, HydroBaseFunctor.h:76-102 (the sound speed is returned
// separately because only the dt reduction needs it).  SHARED = false keeps
// the three IEEE divisions for FP64 as well (fewer registers: the split
// step's FP64 fallback for kinetic-dominated cells).
template <class T, bool SHARED = true>
__device__ __forceinline__ void prim_no_c(const SchemeT<T> &s, T u0, T u1, T u2, T u3,
                                          Q4T<T> &q) {
  using O = Ops<T>;
  if constexpr (SHARED && sizeof(T) == sizeof(double)) {
    prim_fp64(s, u0, u1, u2, u3, q);
  } else {
    const T d = O::mx(u0, s.smallr);
    const T ux = O::div(u2, d);
    const T uy = O::div(u3, d);
    const T eken = T(0.5) * O::mad(uy, uy, ux * ux);
    const T e = O::div(u1, d) - eken;
    const T p = O::mx((s.gamma0 - T(1.0)) * d * e, d * s.smallp);
    q.v[ID] = d;
    q.v[IP] = p;
    q.v[IU] = ux;
    q.v[IV] = uy;
  }
}

// ComputeDtFunctor body, HydroRunFunctors.h:53-74: (c+|u|)/dx + (c+|v|)/dy
template <class T>
__device__ __forceinline__ T cfl_rate_q(const SchemeT<T> &s, const Q4T<T> &q, T dx, T dy) {
  using O = Ops<T>;
  const T c = O::sqroot(O::div(s.gamma0 * (q.v[IP]), q.v[ID]));
  const T vx = c + O::ab(q.v[IU]);
  const T vy = c + O::ab(q.v[IV]);
  return O::div(vx, dx) + O::div(vy, dy);
}

// Cell sizes for the CFL rate.  When a cell size is an exact power of two its
// reciprocal is exact and x / dx == x * (1/dx) for every x (IEEE division by
// 2^-k is an exact scaling, overflow and NaN included), so the division is
// replaced by a multiplication; other sizes keep the division.
template <class T>
struct RateT {
  T dx, dy, idx, idy;
  int pow2x, pow2y;
};

template <class T>
__device__ __forceinline__ T cfl_rate_q(const SchemeT<T> &s, const Q4T<T> &q, const RateT<T> &r) {
  using O = Ops<T>;
  const T c = O::sqroot(O::div(s.gamma0 * (q.v[IP]), q.v[ID]));
  const T vx = c + O::ab(q.v[IU]);
  const T vy = c + O::ab(q.v[IV]);
  T rx, ry;
  if (r.pow2x)  // uniform
    rx = vx * r.idx;
  else
    rx = O::div(vx, r.dx);
  if (r.pow2y)
    ry = vy * r.idy;
  else
    ry = O::div(vy, r.dy);
  return rx + ry;
}

// The rate's square root of gamma p / d.  p >= d smallp, and with smallc^2 >=
// 2^-759 and smallr smallp >= 2^-1000 (the host's floor_sqrt condition, which
// keeps d smallp normal for every d >= smallr) the argument is at least
// smallc^2 (1 - 5u) > 2^-767, where the short sequence gives the IEEE bits;
// +inf keeps sqrt's result, and NaN stays NaN.
template <bool FS>
__device__ __forceinline__ double rate_sqrt(double x) {
  if constexpr (FS) {
    constexpr int kClassPosInf = 1 << 9;
    double r = sqrt_core(x);
    // formed in every lane and then selected: a branch around the root would
    // split the march's row for the lanes holding +inf
    asm volatile("" : "+v"(r));
    return __builtin_amdgcn_class(x, kClassPosInf) ? x : r;
  } else {
    return sqrt(x);
  }
}

// cfl_rate_q for FP64 from the conserved state: the quotient gamma p / d
// shares the primitive conversion's reciprocal and its one wave-level check.
// POW2: both cell sizes are known to be powers of two (no test per cell).
template <bool FS, bool POW2 = false>
__device__ __forceinline__ double cfl_rate_fp64(const Scheme &s, double u0, double u1, double u2,
                                                double u3, const RateT<double> &rt) {
  Q4 q;
  double r;
  bool rok;
  bool ok = prim_fp64_shared(s, u0, u1, u2, u3, q, r, rok);
  double x = quot_shared(s.gamma0 * q.v[IP], q.v[ID], r, rok);
  if (!wave_all(ok && rok)) {  // uniform
    prim_no_c<double, false>(s, u0, u1, u2, u3, q);
    x = (s.gamma0 * q.v[IP]) / q.v[ID];
  }
  const double c = rate_sqrt<FS>(x);
  const double vx = c + fabs(q.v[IU]);
  const double vy = c + fabs(q.v[IV]);
  if constexpr (POW2) return vx * rt.idx + vy * rt.idy;
  double rx, ry;
  if (rt.pow2x)  // uniform
    rx = vx * rt.idx;
  else
    rx = vx / rt.dx;
  if (rt.pow2y)
    ry = vy * rt.idy;
  else
    ry = vy / rt.dy;
  return rx + ry;
}

__device__ __forceinline__ double cfl_rate(const Scheme &s, double u0, double u1, double u2,
                                           double u3, double dx, double dy) {
  Q4 q;
  prim_no_c(s, u0, u1, u2, u3, q);
  return cfl_rate_q(s, q, dx, dy);
}

// slope_unsplit_hydro_2d_scalar, HydroBaseFunctor.h:419-455 (one direction).
// MINMOD: st == 1, whose products with the one-sided differences are exact
// identities (every value, zeros, infinities and NaN included) and are left out.
template <bool MINMOD = false, class T>
__device__ __forceinline__ T limited_slope(T st, T q, T qplus, T qminus) {
  using O = Ops<T>;
  const T dlft = MINMOD ? q - qminus : st * (q - qminus);
  const T drgt = MINMOD ? qplus - q : st * (qplus - q);
  const T dcen = T(0.5) * (qplus - qminus);
  const T dsgn = (dcen >= T(0.0)) ? T(1.0) : T(-1.0);
  const T slop = O::mn(O::ab(dlft), O::ab(drgt));
  const T dlim = ((dlft * drgt) <= T(0.0)) ? T(0.0) : slop;
  return dsgn * O::mn(dlim, O::ab(dcen));
}

// The limited slopes of the four variables of a cell in one direction, from
// its neighbours qp and qm, for slope_type == 1 (MINMOD) or any other non-zero
// slope_type
template <bool MINMOD, class T>
__device__ __forceinline__ void slopes_dir(T st, const Q4T<T> &q, const Q4T<T> &qp,
                                           const Q4T<T> &qm, Q4T<T> &dq) {
  for (int v = 0; v < 4; v++) dq.v[v] = limited_slope<MINMOD>(st, q.v[v], qp.v[v], qm.v[v]);
}

// Both directions, variable by variable (the order the split step's register
// allocation was tuned with)
template <bool MINMOD, class T>
__device__ __forceinline__ void slopes_xy(T st, const Q4T<T> &q, const Q4T<T> &qpx,
                                          const Q4T<T> &qmx, const Q4T<T> &qpy,
                                          const Q4T<T> &qmy, Q4T<T> &dqX, Q4T<T> &dqY) {
  for (int v = 0; v < 4; v++) {
    dqX.v[v] = limited_slope<MINMOD>(st, q.v[v], qpx.v[v], qmx.v[v]);
    dqY.v[v] = limited_slope<MINMOD>(st, q.v[v], qpy.v[v], qmy.v[v]);
  }
}

// slope_unsplit_hydro_2d, HydroBaseFunctor.h:473-516 for one variable: zero
// slopes when slope_type == 0, limited slopes when 1 (minmod) or 2 (MC)
template <class T>
__device__ __forceinline__ T slope1(const SchemeT<T> &s, T q, T qplus, T qminus) {
  // slope_type is uniform, so this is a scalar branch, not VALU work
  return s.slope_type == T(0.0) ? T(0.0) : limited_slope(s.slope_type, q, qplus, qminus);
}

// Source terms of trace_unsplit_2d_along_dir, HydroBaseFunctor.h:246-249.  They
// do not depend on the face, so one evaluation serves the four faces of a cell.
template <class T>
struct SourcesT {
  T sr0, sp0, su0, sv0;
};
using Sources = SourcesT<double>;

template <class T>
__device__ __forceinline__ SourcesT<T> trace_sources(const SchemeT<T> &s, const Q4T<T> &q,
                                                     const Q4T<T> &dqX, const Q4T<T> &dqY) {
  using O = Ops<T>;
  const T r = q.v[ID], p = q.v[IP], u = q.v[IU], v = q.v[IV];
  SourcesT<T> o;
  // ((-u*a - v*b) - c*d) per source term, with the products fused in FP32
  const T divu = dqX.v[IU] + dqY.v[IV];
  o.sr0 = O::mad(-divu, r, O::mad(-v, dqY.v[ID], -u * dqX.v[ID]));
  o.sp0 = O::mad(-(divu * s.gamma0), p, O::mad(-v, dqY.v[IP], -u * dqX.v[IP]));
  o.su0 = O::mad(-v, dqY.v[IU], -u * dqX.v[IU]) - O::div((dqX.v[IP]), r);
  o.sv0 = O::mad(-v, dqY.v[IV], -u * dqX.v[IV]) - O::div((dqY.v[IP]), r);
  return o;
}

// trace_sources for FP64 with the two pressure-slope quotients formed from the
// reciprocal rinv of the density q.v[ID] that the primitive conversion made
// (rok: d and rinv normal), with no fallback; ok is cleared in a lane whose
// quotients may differ from the IEEE ones.
__device__ __forceinline__ Sources trace_sources_shared(const Scheme &s, const Q4 &q,
                                                        const Q4 &dqX, const Q4 &dqY, double rinv,
                                                        bool rok, bool &ok) {
  using O = Ops<double>;
  const double r = q.v[ID], p = q.v[IP], u = q.v[IU], v = q.v[IV];
  Sources o;
  const double divu = dqX.v[IU] + dqY.v[IV];
  o.sr0 = O::mad(-divu, r, O::mad(-v, dqY.v[ID], -u * dqX.v[ID]));
  o.sp0 = O::mad(-(divu * s.gamma0), p, O::mad(-v, dqY.v[IP], -u * dqX.v[IP]));
  bool qok = rok;
  const double qx = quot_shared(dqX.v[IP], r, rinv, qok);
  const double qy = quot_shared(dqY.v[IP], r, rinv, qok);
  ok = ok && qok;
  o.su0 = O::mad(-v, dqY.v[IU], -u * dqX.v[IU]) - qx;
  o.sv0 = O::mad(-v, dqY.v[IV], -u * dqX.v[IV]) - qy;
  return o;
}

// Face states of trace_unsplit_2d_along_dir, HydroBaseFunctor.h:251-289.
// plus = false for the XMIN / YMIN faces, true for XMAX / YMAX; dq is the
// slope along the face normal, dtd = dtdx or dtdy.
template <class T>
__device__ __forceinline__ void trace_face(const SchemeT<T> &s, const Q4T<T> &q, const Q4T<T> &dq,
                                           const SourcesT<T> &src, T dtd, bool plus, Q4T<T> &f) {
  using O = Ops<T>;
  // (q +- h*dq) + (src*dtd)*h
  const T h = T(0.5), hs = plus ? h : -h;
  f.v[ID] = O::mad(src.sr0 * dtd, h, O::mad(hs, dq.v[ID], q.v[ID]));
  f.v[IP] = O::mad(src.sp0 * dtd, h, O::mad(hs, dq.v[IP], q.v[IP]));
  f.v[IU] = O::mad(src.su0 * dtd, h, O::mad(hs, dq.v[IU], q.v[IU]));
  f.v[IV] = O::mad(src.sv0 * dtd, h, O::mad(hs, dq.v[IV], q.v[IV]));
  f.v[ID] = O::mx(s.smallr, f.v[ID]);
}

// riemann_hllc, HydroBaseFunctor.h:704-809.  `ln`/`tn` are the indices of the
// normal / transverse velocity in the Q4 (IU/IV for x faces, IV/IU for y faces,
// which is the reference's swap before and after the call, :628-632).
// The reference evaluates both star states and then picks one; here the side
// (left if ustar > 0) is picked first and its star state evaluated once with
// the same expressions, which gives identical bits for the state it keeps.
template <bool FS = false, class T>
__device__ __forceinline__ void hllc(const SchemeT<T> &s, const Q4T<T> &ql, const Q4T<T> &qr,
                                     int ln, int tn, T flux[4]) {
  using O = Ops<T>;
  const T h = T(0.5);
  const T rl = O::mx(ql.v[ID], s.smallr);
  const T pl = O::mx(ql.v[IP], rl * s.smallp);
  const T ul = ql.v[ln];
  T ecinl = h * rl * ul * ul;
  ecinl = O::mad(h * rl * ql.v[tn], ql.v[tn], ecinl);
  const T etotl = O::mad(pl, s.entho, ecinl);
  const T ptotl = pl;

  const T rr = O::mx(qr.v[ID], s.smallr);
  const T pr = O::mx(qr.v[IP], rr * s.smallp);
  const T ur = qr.v[ln];
  T ecinr = h * rr * ur * ur;
  ecinr = O::mad(h * rr * qr.v[tn], qr.v[tn], ecinr);
  const T etotr = O::mad(pr, s.entho, ecinr);
  const T ptotr = pr;

  const T cfastl = floored_sqrt<FS>(O::mx(O::div(s.gamma0 * pl, rl), s.smallc2));
  const T cfastr = floored_sqrt<FS>(O::mx(O::div(s.gamma0 * pr, rr), s.smallc2));
  const T cmax = O::mx(cfastl, cfastr);
  const T SL = O::mn(ul, ur) - cmax;
  const T SR = O::mx(ul, ur) + cmax;

  const T rcl = rl * (ul - SL);
  const T rcr = rr * (SR - ur);
  // separate products: for mirrored states at a reflecting wall they cancel
  // exactly, so ustar and the wall's mass flux are exactly zero
  const T ustar = O::div(rcr * ur + rcl * ul + (ptotl - ptotr), rcr + rcl);
  // the left/right-symmetric sum stays unfused, so that mirrored states give
  // mirrored fluxes in FP32 too
  const T ptotstar = O::div(O::mad(rcl * rcr, ul - ur, rcr * ptotl + rcl * ptotr), rcr + rcl);

  const bool left = ustar > T(0.0);
  const T rk = left ? rl : rr;
  const T uk = left ? ul : ur;
  const T Sk = left ? SL : SR;
  const T ek = left ? etotl : etotr;
  const T pk = left ? ptotl : ptotr;
  const T rstar = O::div(rk * (Sk - uk), Sk - ustar);
  const T etotstar = O::div(O::mad(ptotstar, ustar, O::mad(-pk, uk, (Sk - uk) * ek)), Sk - ustar);

  T ro, uo, ptoto, etoto;
  if (SL > T(0.0)) {
    ro = rl; uo = ul; ptoto = ptotl; etoto = etotl;
  } else if (left || SR > T(0.0)) {
    ro = rstar; uo = ustar; ptoto = ptotstar; etoto = etotstar;
  } else {
    ro = rr; uo = ur; ptoto = ptotr; etoto = etotr;
  }
  const T fd = ro * uo;
  flux[ID] = fd;
  flux[ln] = O::mad(ro * uo, uo, ptoto);
  flux[IP] = (etoto + ptoto) * uo;
  flux[tn] = fd * ((fd > T(0.0)) ? ql.v[tn] : qr.v[tn]);
}

// HLL (no reference counterpart; the reference declares the enum only,
// HydroParams.h:49).  Toro 2009, section 10.3, eq. (10.21), with the floors and
// wave speeds of riemann_hllc.  Same expressions as the CPU oracle.
template <bool FS = false, class T>
__device__ __forceinline__ void hll(const SchemeT<T> &s, const Q4T<T> &ql, const Q4T<T> &qr,
                                    int ln, int tn, T flux[4]) {
  using O = Ops<T>;
  const T h = T(0.5);
  const T rl = O::mx(ql.v[ID], s.smallr);
  const T pl = O::mx(ql.v[IP], rl * s.smallp);
  const T ul = ql.v[ln], vl = ql.v[tn];
  const T rr = O::mx(qr.v[ID], s.smallr);
  const T pr = O::mx(qr.v[IP], rr * s.smallp);
  const T ur = qr.v[ln], vr = qr.v[tn];
  const T el = O::mad(pl, s.entho, O::mad(h * rl * vl, vl, h * rl * ul * ul));
  const T er = O::mad(pr, s.entho, O::mad(h * rr * vr, vr, h * rr * ur * ur));
  const T cfastl = floored_sqrt<FS>(O::mx(O::div(s.gamma0 * pl, rl), s.smallc2));
  const T cfastr = floored_sqrt<FS>(O::mx(O::div(s.gamma0 * pr, rr), s.smallc2));
  const T SL = O::mn(ul, ur) - O::mx(cfastl, cfastr);
  const T SR = O::mx(ul, ur) + O::mx(cfastl, cfastr);
  // variable order of the conserved state: rho, E, normal momentum, transverse momentum
  const T UL[4] = {rl, el, rl * ul, rl * vl};
  const T UR[4] = {rr, er, rr * ur, rr * vr};
  const T FL[4] = {rl * ul, (el + pl) * ul, O::mad(rl * ul, ul, pl), rl * ul * vl};
  const T FR[4] = {rr * ur, (er + pr) * ur, O::mad(rr * ur, ur, pr), rr * ur * vr};
  T f[4];
  if (SL >= T(0.0)) {
    for (int k = 0; k < 4; k++) f[k] = FL[k];
  } else if (SR <= T(0.0)) {
    for (int k = 0; k < 4; k++) f[k] = FR[k];
  } else {
    const T inv = O::div(T(1.0), SR - SL);
    // SR*FL - SL*FR with separate products: exactly zero mass flux for
    // mirrored states at a reflecting wall
    for (int k = 0; k < 4; k++)
      f[k] = O::mad(SL * SR, UR[k] - UL[k], SR * FL[k] - SL * FR[k]) * inv;
  }
  flux[ID] = f[0];
  flux[IP] = f[1];
  flux[ln] = f[2];
  flux[tn] = f[3];
}

// riemann_approx + cmpflx, HydroBaseFunctor.h:523-693 (ten Newton iterations at
// most, hard-coded at :592)
template <class T>
__device__ __forceinline__ void approx(const SchemeT<T> &s, const Q4T<T> &qlf, const Q4T<T> &qrf,
                                       int ln, int tn, T flux[4]) {
  using O = Ops<T>;
  const T one = T(1.0), h = T(0.5), zero = T(0.0);
  const T rl = O::mx(qlf.v[ID], s.smallr);
  const T ul = qlf.v[ln];
  const T pl = O::mx(qlf.v[IP], rl * s.smallp);
  const T rr = O::mx(qrf.v[ID], s.smallr);
  const T ur = qrf.v[ln];
  const T pr = O::mx(qrf.v[IP], rr * s.smallp);
  const T cl = s.gamma0 * pl * rl;
  const T cr = s.gamma0 * pr * rr;
  T wl = O::sqroot(cl);
  T wr = O::sqroot(cr);
  T pstar = O::mx(O::div((wr * pl + wl * pr) + wl * wr * (ul - ur), wl + wr), zero);
  T pold = pstar;
  T conv = one;
  for (int iter = 0; (iter < 10) && (conv > T(1e-6)); ++iter) {
    const T wwl = O::sqroot(cl * (one + O::div(s.gamma6 * (pold - pl), pl)));
    const T wwr = O::sqroot(cr * (one + O::div(s.gamma6 * (pold - pr), pr)));
    const T qwl = O::div(T(2.0) * wwl * wwl * wwl, wwl * wwl + cl);
    const T qwr = O::div(T(2.0) * wwr * wwr * wwr, wwr * wwr + cr);
    const T usl = ul - O::div(pold - pl, wwl);
    const T usr = ur + O::div(pold - pr, wwr);
    const T delp = O::mx(O::div(qwr * qwl, qwr + qwl) * (usl - usr), -pold);
    pold = pold + delp;
    conv = O::ab(O::div(delp, pold + s.smallpp));
  }
  pstar = pold;
  wl = O::sqroot(cl * (one + O::div(s.gamma6 * (pstar - pl), pl)));
  wr = O::sqroot(cr * (one + O::div(s.gamma6 * (pstar - pr), pr)));
  const T ustar = h * (ul + O::div(pl - pstar, wl) + ur - O::div(pr - pstar, wr));
  const T sgnm = copysign(one, ustar);
  const bool lft = sgnm > zero;
  const T ro = lft ? rl : rr, uo = lft ? ul : ur, po = lft ? pl : pr, wo = lft ? wl : wr;
  const T co = O::mx(s.smallc, O::sqroot(O::ab(O::div(s.gamma0 * po, ro))));
  const T rstar = O::mx((O::div(ro, one + O::div(ro * (po - pstar), wo * wo))), s.smallr);
  const T cstar = O::mx(s.smallc, O::sqroot(O::ab(O::div(s.gamma0 * pstar, rstar))));
  T spout = co - sgnm * uo;
  T spin = cstar - sgnm * ustar;
  const T ushock = O::div(wo, ro) - sgnm * uo;
  if (pstar >= po) {
    spin = ushock;
    spout = ushock;
  }
  const T scr = O::mx(spout - spin, s.smallc + O::ab(spout + spin));
  T frac = h * (one + O::div(spout + spin, scr));
  if (frac != frac)
    frac = zero;
  else
    frac = frac >= one ? one : frac <= zero ? zero : frac;
  T gd = frac * rstar + (one - frac) * ro;
  T gu = frac * ustar + (one - frac) * uo;
  T gp = frac * pstar + (one - frac) * po;
  if (spout < zero) {
    gd = ro;
    gu = uo;
    gp = po;
  }
  if (spin > zero) {
    gd = rstar;
    gu = ustar;
    gp = pstar;
  }
  const T gv = lft ? qlf.v[tn] : qrf.v[tn];
  // cmpflx, HydroBaseFunctor.h:523-547
  const T fd = gd * gu;
  flux[ID] = fd;
  flux[ln] = fd * gu + gp;
  flux[tn] = fd * gv;
  const T ekin = h * gd * (gu * gu + gv * gv);
  const T etot = gp * s.entho + ekin;
  flux[IP] = gu * (etot + gp);
}

// solver chosen at compile time: RS = 0 approx, 1 hll, 2 hllc
// RS: 0 approx, 1 HLL, 2 HLLC; + kRiemannFloorSqrt: the HLL/HLLC wave speeds
// use the short square root (FP64, arguments floored at smallc^2 >= 2^-759)
constexpr int kRiemannFloorSqrt = 8;
// + kMinmodPow2: the march is built for slope_type == 1 (its slopes) and for
// cell sizes that are powers of two (its CFL rate); the host selects such a
// kernel only then
constexpr int kMinmodPow2 = 16;
// + kSlopeMinmod: the march's slopes alone are built for slope_type == 1 (the
// batch kernels, whose instances' cell sizes may be anything)
constexpr int kSlopeMinmod = 32;

template <int RS, class T>
__device__ __forceinline__ void riemann(const SchemeT<T> &s, const Q4T<T> &ql, const Q4T<T> &qr,
                                        int ln, int tn, T flux[4]) {
  constexpr int kind = RS & 7;
  constexpr bool fs = (RS & kRiemannFloorSqrt) != 0;
  if constexpr (kind == 2)
    hllc<fs>(s, ql, qr, ln, tn, flux);
  else if constexpr (kind == 1)
    hll<fs>(s, ql, qr, ln, tn, flux);
  else
    approx(s, ql, qr, ln, tn, flux);
}

}  // namespace e2d
