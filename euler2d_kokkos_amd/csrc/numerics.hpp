// numerics.hpp -- per-cell / per-face arithmetic of the MUSCL-Hancock scheme.
//
// Each function evaluates the same IEEE operations, in the same order, as the
// reference's inline functions (pkestene/euler2d_kokkos src/HydroBaseFunctor.h),
// so that the FP64 path reproduces the reference bit for bit: this file must be
// compiled without FMA contraction (-ffp-contract=off and the pragma below) and
// without fast-math.  Restructurings that keep every result bit-identical are
// marked where they happen (shared source terms, one star state per HLLC).
#pragma once

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace e2d {

enum { ID = 0, IP = 1, IU = 2, IV = 3 };

// Scheme constants, precomputed on the host in double (entho = 1/(gamma0-1) is
// the same IEEE division the reference performs inside every Riemann call).
struct Scheme {
  double gamma0, gamma6, smallr, smallc, smallp, smallpp, slope_type, entho;
  double smallc2;  // smallc * smallc
  int riemann;     // 0 approx, 1 hll, 2 hllc (already resolved from honor_riemann)
};

struct Q4 {
  double v[4];
};

// computePrimitives, HydroBaseFunctor.h:76-102 (the sound speed is returned
// separately because only the dt reduction needs it)
__device__ __forceinline__ void prim_no_c(const Scheme &s, double u0, double u1, double u2,
                                          double u3, Q4 &q) {
  const double d = fmax(u0, s.smallr);
  const double ux = u2 / d;
  const double uy = u3 / d;
  const double eken = 0.5 * (ux * ux + uy * uy);
  const double e = u1 / d - eken;
  const double p = fmax((s.gamma0 - 1.0) * d * e, d * s.smallp);
  q.v[ID] = d;
  q.v[IP] = p;
  q.v[IU] = ux;
  q.v[IV] = uy;
}

// ComputeDtFunctor body, HydroRunFunctors.h:53-74: (c+|u|)/dx + (c+|v|)/dy
__device__ __forceinline__ double cfl_rate(const Scheme &s, double u0, double u1, double u2,
                                           double u3, double dx, double dy) {
  Q4 q;
  prim_no_c(s, u0, u1, u2, u3, q);
  const double c = sqrt(s.gamma0 * (q.v[IP]) / q.v[ID]);
  const double vx = c + fabs(q.v[IU]);
  const double vy = c + fabs(q.v[IV]);
  return vx / dx + vy / dy;
}

// slope_unsplit_hydro_2d_scalar, HydroBaseFunctor.h:419-455 (one direction)
__device__ __forceinline__ double limited_slope(double st, double q, double qplus, double qminus) {
  const double dlft = st * (q - qminus);
  const double drgt = st * (qplus - q);
  const double dcen = 0.5 * (qplus - qminus);
  const double dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
  const double slop = fmin(fabs(dlft), fabs(drgt));
  const double dlim = ((dlft * drgt) <= 0.0) ? 0.0 : slop;
  return dsgn * fmin(dlim, fabs(dcen));
}

// slope_unsplit_hydro_2d, HydroBaseFunctor.h:473-516 for one variable: zero
// slopes when slope_type == 0, limited slopes when 1 (minmod) or 2 (MC)
__device__ __forceinline__ double slope1(const Scheme &s, double q, double qplus, double qminus) {
  return s.slope_type == 0.0 ? 0.0 : limited_slope(s.slope_type, q, qplus, qminus);
}

// Source terms of trace_unsplit_2d_along_dir, HydroBaseFunctor.h:246-249.  They
// do not depend on the face, so one evaluation serves the four faces of a cell.
struct Sources {
  double sr0, sp0, su0, sv0;
};

__device__ __forceinline__ Sources trace_sources(const Scheme &s, const Q4 &q, const Q4 &dqX,
                                                 const Q4 &dqY) {
  const double r = q.v[ID], p = q.v[IP], u = q.v[IU], v = q.v[IV];
  Sources o;
  o.sr0 = -u * dqX.v[ID] - v * dqY.v[ID] - (dqX.v[IU] + dqY.v[IV]) * r;
  o.sp0 = -u * dqX.v[IP] - v * dqY.v[IP] - (dqX.v[IU] + dqY.v[IV]) * s.gamma0 * p;
  o.su0 = -u * dqX.v[IU] - v * dqY.v[IU] - (dqX.v[IP]) / r;
  o.sv0 = -u * dqX.v[IV] - v * dqY.v[IV] - (dqY.v[IP]) / r;
  return o;
}

// Face states of trace_unsplit_2d_along_dir, HydroBaseFunctor.h:251-289.
// sgn = -1 for the XMIN / YMIN faces, +1 for XMAX / YMAX; dq is the slope
// along the face normal, dtd = dtdx or dtdy.
__device__ __forceinline__ void trace_face(const Scheme &s, const Q4 &q, const Q4 &dq,
                                           const Sources &src, double dtd, bool plus, Q4 &f) {
  if (plus) {
    f.v[ID] = q.v[ID] + 0.5 * dq.v[ID] + src.sr0 * dtd * 0.5;
    f.v[IP] = q.v[IP] + 0.5 * dq.v[IP] + src.sp0 * dtd * 0.5;
    f.v[IU] = q.v[IU] + 0.5 * dq.v[IU] + src.su0 * dtd * 0.5;
    f.v[IV] = q.v[IV] + 0.5 * dq.v[IV] + src.sv0 * dtd * 0.5;
  } else {
    f.v[ID] = q.v[ID] - 0.5 * dq.v[ID] + src.sr0 * dtd * 0.5;
    f.v[IP] = q.v[IP] - 0.5 * dq.v[IP] + src.sp0 * dtd * 0.5;
    f.v[IU] = q.v[IU] - 0.5 * dq.v[IU] + src.su0 * dtd * 0.5;
    f.v[IV] = q.v[IV] - 0.5 * dq.v[IV] + src.sv0 * dtd * 0.5;
  }
  f.v[ID] = fmax(s.smallr, f.v[ID]);
}

// riemann_hllc, HydroBaseFunctor.h:704-809.  `ln`/`tn` are the indices of the
// normal / transverse velocity in the Q4 (IU/IV for x faces, IV/IU for y faces,
// which is the reference's swap before and after the call, :628-632).
// The reference evaluates both star states and then picks one; here the side
// (left if ustar > 0) is picked first and its star state evaluated once with
// the same expressions, which gives identical bits for the state it keeps.
__device__ __forceinline__ void hllc(const Scheme &s, const Q4 &ql, const Q4 &qr, int ln, int tn,
                                     double flux[4]) {
  const double rl = fmax(ql.v[ID], s.smallr);
  const double pl = fmax(ql.v[IP], rl * s.smallp);
  const double ul = ql.v[ln];
  double ecinl = 0.5 * rl * ul * ul;
  ecinl += 0.5 * rl * ql.v[tn] * ql.v[tn];
  const double etotl = pl * s.entho + ecinl;
  const double ptotl = pl;

  const double rr = fmax(qr.v[ID], s.smallr);
  const double pr = fmax(qr.v[IP], rr * s.smallp);
  const double ur = qr.v[ln];
  double ecinr = 0.5 * rr * ur * ur;
  ecinr += 0.5 * rr * qr.v[tn] * qr.v[tn];
  const double etotr = pr * s.entho + ecinr;
  const double ptotr = pr;

  const double cfastl = sqrt(fmax(s.gamma0 * pl / rl, s.smallc2));
  const double cfastr = sqrt(fmax(s.gamma0 * pr / rr, s.smallc2));
  const double cmax = fmax(cfastl, cfastr);
  const double SL = fmin(ul, ur) - cmax;
  const double SR = fmax(ul, ur) + cmax;

  const double rcl = rl * (ul - SL);
  const double rcr = rr * (SR - ur);
  const double ustar = (rcr * ur + rcl * ul + (ptotl - ptotr)) / (rcr + rcl);
  const double ptotstar = (rcr * ptotl + rcl * ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

  const bool left = ustar > 0.0;
  const double rk = left ? rl : rr;
  const double uk = left ? ul : ur;
  const double Sk = left ? SL : SR;
  const double ek = left ? etotl : etotr;
  const double pk = left ? ptotl : ptotr;
  const double rstar = rk * (Sk - uk) / (Sk - ustar);
  const double etotstar = ((Sk - uk) * ek - pk * uk + ptotstar * ustar) / (Sk - ustar);

  double ro, uo, ptoto, etoto;
  if (SL > 0.0) {
    ro = rl; uo = ul; ptoto = ptotl; etoto = etotl;
  } else if (left || SR > 0.0) {
    ro = rstar; uo = ustar; ptoto = ptotstar; etoto = etotstar;
  } else {
    ro = rr; uo = ur; ptoto = ptotr; etoto = etotr;
  }
  const double fd = ro * uo;
  flux[ID] = fd;
  flux[ln] = ro * uo * uo + ptoto;
  flux[IP] = (etoto + ptoto) * uo;
  flux[tn] = fd * ((fd > 0.0) ? ql.v[tn] : qr.v[tn]);
}

// HLL (no reference counterpart; the reference declares the enum only,
// HydroParams.h:49).  Toro 2009, section 10.3, eq. (10.21), with the floors and
// wave speeds of riemann_hllc.  Same expressions as the CPU oracle.
__device__ __forceinline__ void hll(const Scheme &s, const Q4 &ql, const Q4 &qr, int ln, int tn,
                                    double flux[4]) {
  const double rl = fmax(ql.v[ID], s.smallr);
  const double pl = fmax(ql.v[IP], rl * s.smallp);
  const double ul = ql.v[ln], vl = ql.v[tn];
  const double rr = fmax(qr.v[ID], s.smallr);
  const double pr = fmax(qr.v[IP], rr * s.smallp);
  const double ur = qr.v[ln], vr = qr.v[tn];
  const double el = pl * s.entho + (0.5 * rl * ul * ul + 0.5 * rl * vl * vl);
  const double er = pr * s.entho + (0.5 * rr * ur * ur + 0.5 * rr * vr * vr);
  const double cfastl = sqrt(fmax(s.gamma0 * pl / rl, s.smallc2));
  const double cfastr = sqrt(fmax(s.gamma0 * pr / rr, s.smallc2));
  const double SL = fmin(ul, ur) - fmax(cfastl, cfastr);
  const double SR = fmax(ul, ur) + fmax(cfastl, cfastr);
  // variable order of the conserved state: rho, E, normal momentum, transverse momentum
  const double UL[4] = {rl, el, rl * ul, rl * vl};
  const double UR[4] = {rr, er, rr * ur, rr * vr};
  const double FL[4] = {rl * ul, (el + pl) * ul, rl * ul * ul + pl, rl * ul * vl};
  const double FR[4] = {rr * ur, (er + pr) * ur, rr * ur * ur + pr, rr * ur * vr};
  double f[4];
  if (SL >= 0.0) {
    for (int k = 0; k < 4; k++) f[k] = FL[k];
  } else if (SR <= 0.0) {
    for (int k = 0; k < 4; k++) f[k] = FR[k];
  } else {
    const double inv = 1.0 / (SR - SL);
    for (int k = 0; k < 4; k++) f[k] = (SR * FL[k] - SL * FR[k] + SL * SR * (UR[k] - UL[k])) * inv;
  }
  flux[ID] = f[0];
  flux[IP] = f[1];
  flux[ln] = f[2];
  flux[tn] = f[3];
}

// riemann_approx + cmpflx, HydroBaseFunctor.h:523-693 (ten Newton iterations at
// most, hard-coded at :592)
__device__ __forceinline__ void approx(const Scheme &s, const Q4 &qlf, const Q4 &qrf, int ln,
                                       int tn, double flux[4]) {
  const double rl = fmax(qlf.v[ID], s.smallr);
  const double ul = qlf.v[ln];
  const double pl = fmax(qlf.v[IP], rl * s.smallp);
  const double rr = fmax(qrf.v[ID], s.smallr);
  const double ur = qrf.v[ln];
  const double pr = fmax(qrf.v[IP], rr * s.smallp);
  const double cl = s.gamma0 * pl * rl;
  const double cr = s.gamma0 * pr * rr;
  double wl = sqrt(cl);
  double wr = sqrt(cr);
  double pstar = fmax(((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr), 0.0);
  double pold = pstar;
  double conv = 1.0;
  for (int iter = 0; (iter < 10) && (conv > 1e-6); ++iter) {
    const double wwl = sqrt(cl * (1.0 + s.gamma6 * (pold - pl) / pl));
    const double wwr = sqrt(cr * (1.0 + s.gamma6 * (pold - pr) / pr));
    const double qwl = 2.0 * wwl * wwl * wwl / (wwl * wwl + cl);
    const double qwr = 2.0 * wwr * wwr * wwr / (wwr * wwr + cr);
    const double usl = ul - (pold - pl) / wwl;
    const double usr = ur + (pold - pr) / wwr;
    const double delp = fmax(qwr * qwl / (qwr + qwl) * (usl - usr), -pold);
    pold = pold + delp;
    conv = fabs(delp / (pold + s.smallpp));
  }
  pstar = pold;
  wl = sqrt(cl * (1.0 + s.gamma6 * (pstar - pl) / pl));
  wr = sqrt(cr * (1.0 + s.gamma6 * (pstar - pr) / pr));
  const double ustar = 0.5 * (ul + (pl - pstar) / wl + ur - (pr - pstar) / wr);
  const double sgnm = copysign(1.0, ustar);
  const bool lft = sgnm > 0.0;
  const double ro = lft ? rl : rr, uo = lft ? ul : ur, po = lft ? pl : pr, wo = lft ? wl : wr;
  const double co = fmax(s.smallc, sqrt(fabs(s.gamma0 * po / ro)));
  const double rstar = fmax((ro / (1.0 + ro * (po - pstar) / (wo * wo))), s.smallr);
  const double cstar = fmax(s.smallc, sqrt(fabs(s.gamma0 * pstar / rstar)));
  double spout = co - sgnm * uo;
  double spin = cstar - sgnm * ustar;
  const double ushock = wo / ro - sgnm * uo;
  if (pstar >= po) {
    spin = ushock;
    spout = ushock;
  }
  const double scr = fmax(spout - spin, s.smallc + fabs(spout + spin));
  double frac = 0.5 * (1.0 + (spout + spin) / scr);
  if (frac != frac)
    frac = 0.0;
  else
    frac = frac >= 1.0 ? 1.0 : frac <= 0.0 ? 0.0 : frac;
  double gd = frac * rstar + (1.0 - frac) * ro;
  double gu = frac * ustar + (1.0 - frac) * uo;
  double gp = frac * pstar + (1.0 - frac) * po;
  if (spout < 0.0) {
    gd = ro;
    gu = uo;
    gp = po;
  }
  if (spin > 0.0) {
    gd = rstar;
    gu = ustar;
    gp = pstar;
  }
  const double gv = lft ? qlf.v[tn] : qrf.v[tn];
  // cmpflx, HydroBaseFunctor.h:523-547
  const double fd = gd * gu;
  flux[ID] = fd;
  flux[ln] = fd * gu + gp;
  flux[tn] = fd * gv;
  const double ekin = 0.5 * gd * (gu * gu + gv * gv);
  const double etot = gp * s.entho + ekin;
  flux[IP] = gu * (etot + gp);
}

// solver chosen at compile time: RS = 0 approx, 1 hll, 2 hllc
template <int RS>
__device__ __forceinline__ void riemann(const Scheme &s, const Q4 &ql, const Q4 &qr, int ln, int tn,
                                        double flux[4]) {
  if constexpr (RS == 2)
    hllc(s, ql, qr, ln, tn, flux);
  else if constexpr (RS == 1)
    hll(s, ql, qr, ln, tn, flux);
  else
    approx(s, ql, qr, ln, tn, flux);
}

}  // namespace e2d
