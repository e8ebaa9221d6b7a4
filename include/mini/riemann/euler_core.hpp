// Euler physics of the MI355X kernels in portable form (FP64): ideal gas,
// the rotated frame, the Exact / HLLC / Roe upwind fluxes and the boundary
// fluxes.  The device kernels compile it with hipcc; the host C++ surface
// (mini/riemann/euler.hpp) compiles the same code with any C++17 compiler,
// so a host caller gets the arithmetic the kernels run.
//
// Reference arithmetic:
//   ideal gas, flux matrix   riemann/euler/types.hpp:186-334 (gamma = 1.4)
//   rotated frame, BC fluxes riemann/rotated/euler.hpp:117-174,181-280,296-299,340-345
//   exact solver             riemann/euler/exact.hpp:62-320 (Toro ch. 4)
//   HLLC                     riemann/euler/hllc.hpp:30-109 (Toro ch. 10)
//   Roe                      python/riemann.py:414-451 (1-D, no entropy fix),
//                            extended here with two shear waves (Toro 11.2)
// States are five doubles (rho, rho u, rho v, rho w, E); primitive states are
// (rho, u, v, w, p).  Every function is branch-light straight-line code so a
// wavefront of face points stays converged except inside the exact solver.
#ifndef MINI_RIEMANN_EULER_CORE_HPP_
#define MINI_RIEMANN_EULER_CORE_HPP_

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MCFD_HD __host__ __device__ __forceinline__
#else
#define MCFD_HD inline
#endif

namespace mcfd {

constexpr double kGamma = 1.4;
constexpr double kGm1 = kGamma - 1;
constexpr double kGp1 = kGamma + 1;

struct State { double v[5]; };

// types.hpp:195-202,289-301: conservative -> primitive, zeroing the whole
// state when rho < 0 or p < 0.
MCFD_HD State ToPrimitive(State const &c) {
  State p = c;
  if (p.v[0] < 0 || p.v[4] < 0) {
    for (int i = 0; i < 5; ++i) p.v[i] = 0;
  }
  if (p.v[0] != 0) {
    double rho = p.v[0];
    p.v[1] /= rho; p.v[2] /= rho; p.v[3] /= rho;
    double ke = 0.5 * (p.v[1] * p.v[1] + p.v[2] * p.v[2] + p.v[3] * p.v[3]);
    p.v[4] -= rho * ke;
    p.v[4] *= kGm1;
    if (p.v[0] < 0 || p.v[4] < 0) {
      for (int i = 0; i < 5; ++i) p.v[i] = 0;
    }
  }
  return p;
}

MCFD_HD State ToConservative(State const &p) {  // types.hpp:303-311
  State c = p;
  c.v[1] *= p.v[0]; c.v[2] *= p.v[0]; c.v[3] *= p.v[0];
  c.v[4] *= 1 / kGm1;
  c.v[4] += p.v[0] * (0.5 * (p.v[1] * p.v[1] + p.v[2] * p.v[2] + p.v[3] * p.v[3]));
  return c;
}

MCFD_HD State NormalFlux(State const &p) {  // types.hpp:313-321, normal frame
  State f = ToConservative(p);
  for (int i = 0; i < 5; ++i) f.v[i] *= p.v[1];
  f.v[1] += p.v[4];
  f.v[4] += p.v[4] * p.v[1];
  return f;
}

MCFD_HD double SoundSpeed(double rho, double p) {  // types.hpp:252-254
  return (rho > 0 && p > 0) ? sqrt(kGamma * p / rho) : 0;
}

// types.hpp:322-333 physical flux, f[k][d] for k in 0..4, d in 0..2.
MCFD_HD void FluxMatrix(State const &c, double f[5][3]) {
  State p = ToPrimitive(c);
  double u = p.v[1], v = p.v[2], w = p.v[3], pr = p.v[4];
  double ru = c.v[1], rv = c.v[2], rw = c.v[3], rh = c.v[4] + pr;
  f[0][0] = ru;          f[0][1] = rv;          f[0][2] = rw;
  f[1][0] = ru * u + pr; f[1][1] = ru * v;      f[1][2] = ru * w;
  f[2][0] = rv * u;      f[2][1] = rv * v + pr; f[2][2] = rv * w;
  f[3][0] = rw * u;      f[3][1] = rw * v;      f[3][2] = rw * w + pr;
  f[4][0] = rh * u;      f[4][1] = rh * v;      f[4][2] = rh * w;
}

// rotated/euler.hpp:340-345
MCFD_HD double MaxSpeed(State const &c) {
  State p = ToPrimitive(c);
  double a = SoundSpeed(p.v[0], p.v[4]);
  double u = sqrt(p.v[1] * p.v[1] + p.v[2] * p.v[2] + p.v[3] * p.v[3]);
  return fmax(a + u, a - u);
}

struct FaceFrame { double n[3], t[3], b[3]; };

MCFD_HD void GlobalToNormal(FaceFrame const &f, State *s) {  // euler.hpp:117-125
  double x = s->v[1], y = s->v[2], z = s->v[3];
  s->v[1] = x * f.n[0] + y * f.n[1] + z * f.n[2];
  s->v[2] = x * f.t[0] + y * f.t[1] + z * f.t[2];
  s->v[3] = x * f.b[0] + y * f.b[1] + z * f.b[2];
}
MCFD_HD void NormalToGlobal(FaceFrame const &f, State *s) {  // euler.hpp:126-134
  double a = s->v[1], b = s->v[2], c = s->v[3];
  s->v[1] = a * f.n[0] + b * f.t[0] + c * f.b[0];
  s->v[2] = a * f.n[1] + b * f.t[1] + c * f.b[1];
  s->v[3] = a * f.n[2] + b * f.t[2] + c * f.b[2];
}

// ------------------------------------------------------------------ HLLC
MCFD_HD double HllcQ(double p_est, double pk) {  // hllc.hpp:70-78
  if (p_est <= pk) return 1.0;
  return sqrt(1 + (kGp1 / 2) * (p_est / pk - 1) / kGamma);
}
MCFD_HD State HllcStar(State const &s, double wk, double ws) {  // hllc.hpp:94-109
  State flux = NormalFlux(s);
  State uk = ToConservative(s);
  double temp = s.v[0] * (wk - s.v[1]) / (wk - ws);
  State us;
  us.v[0] = temp;
  us.v[1] = temp * ws; us.v[2] = temp * s.v[2]; us.v[3] = temp * s.v[3];
  us.v[4] = uk.v[4] / s.v[0] + (ws - s.v[1]) * (ws + s.v[4] / (s.v[0] * (wk - s.v[1])));
  us.v[4] *= temp;
  for (int i = 0; i < 5; ++i) flux.v[i] += (us.v[i] - uk.v[i]) * wk;
  return flux;
}
MCFD_HD State HllcFlux(State const &l, State const &r) {  // hllc.hpp:30-69
  double rho_avg = (l.v[0] + r.v[0]) / 2;
  double al = SoundSpeed(l.v[0], l.v[4]), ar = SoundSpeed(r.v[0], r.v[4]);
  double a_avg = (al + ar) / 2;
  double p_pvrs = (l.v[4] + r.v[4] - (r.v[1] - l.v[1]) * rho_avg * a_avg) / 2;
  double p_est = fmax(0.0, p_pvrs);
  double wl = l.v[1] - al * HllcQ(p_est, l.v[4]);
  double wr = r.v[1] + ar * HllcQ(p_est, r.v[4]);
  double ws = (r.v[4] - l.v[4] + l.v[0] * l.v[1] * (wl - l.v[1]) - r.v[0] * r.v[1] * (wr - r.v[1]))
            / (l.v[0] * (wl - l.v[1]) - r.v[0] * (wr - r.v[1]));
  if (0.0 <= wl) return NormalFlux(l);
  if (wr <= 0.0) return NormalFlux(r);
  if (0.0 <= ws) return HllcStar(l, wl, ws);
  return HllcStar(r, wr, ws);
}

// ------------------------------------------------------------------- Roe
MCFD_HD State RoeFlux(State const &l, State const &r) {
  State fl = NormalFlux(l), fr = NormalFlux(r), flux;
  for (int i = 0; i < 5; ++i) flux.v[i] = (fl.v[i] + fr.v[i]) / 2;
  constexpr double kGoGm1 = kGamma / kGm1;
  double h0l = l.v[4] / l.v[0] * kGoGm1 + 0.5 * (l.v[1] * l.v[1] + l.v[2] * l.v[2] + l.v[3] * l.v[3]);
  double h0r = r.v[4] / r.v[0] * kGoGm1 + 0.5 * (r.v[1] * r.v[1] + r.v[2] * r.v[2] + r.v[3] * r.v[3]);
  double sl = sqrt(l.v[0]), sr = sqrt(r.v[0]), inv = 1 / (sl + sr);
  double rho = sl * sr;
  double u = (sl * l.v[1] + sr * r.v[1]) * inv;
  double v = (sl * l.v[2] + sr * r.v[2]) * inv;
  double w = (sl * l.v[3] + sr * r.v[3]) * inv;
  double h = (sl * h0l + sr * h0r) * inv;
  double ke = (u * u + v * v + w * w) / 2;
  double aa = kGm1 * (h - ke);
  double a = sqrt(aa), ua = u * a;
  double dp = r.v[4] - l.v[4], du = r.v[1] - l.v[1];
  double s;
  s = 0.5 * ((dp - rho * a * du) / (2 * aa)) * fabs(u - a);  // u - a
  flux.v[0] -= s; flux.v[1] -= s * (u - a); flux.v[2] -= s * v; flux.v[3] -= s * w; flux.v[4] -= s * (h - ua);
  s = 0.5 * ((r.v[0] - l.v[0]) - dp / aa) * fabs(u);          // entropy
  flux.v[0] -= s; flux.v[1] -= s * u; flux.v[2] -= s * v; flux.v[3] -= s * w; flux.v[4] -= s * ke;
  s = 0.5 * (rho * (r.v[2] - l.v[2])) * fabs(u);                // shear (t)
  flux.v[2] -= s; flux.v[4] -= s * v;
  s = 0.5 * (rho * (r.v[3] - l.v[3])) * fabs(u);                // shear (b)
  flux.v[3] -= s; flux.v[4] -= s * w;
  s = 0.5 * ((dp + rho * a * du) / (2 * aa)) * fabs(u + a);   // u + a
  flux.v[0] -= s; flux.v[1] -= s * (u + a); flux.v[2] -= s * v; flux.v[3] -= s * w; flux.v[4] -= s * (h + ua);
  return flux;
}

// ------------------------------------------------- unfused rounding
// One IEEE-rounded add or multiply that FP contraction may not merge into an
// FMA with its neighbours: HIP's __dadd_rn / __dmul_rn are plain operators,
// which the default contraction fuses (u + r * dt becomes one FMA).  Used
// where the reference's operation order has to be kept to the last bit: the
// RK combinations of temporal/rk.hpp:31-49 and ode.hpp:121-128, and the
// host-order projection of Approximate.
MCFD_HD double AddRn(double a, double b) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return a + b;
}
MCFD_HD double MulRn(double a, double b) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return a * b;
}

// ----------------------------------------------------------------- exact
struct ExactWave {  // exact.hpp:153-200, the pressure function of one side
  double rho, p, a, p_const;
  MCFD_HD explicit ExactWave(State const &b)
      : rho(b.v[0]), p(b.v[4]), a(b.v[0] == 0 ? 0 : SoundSpeed(b.v[0], b.v[4])),
        p_const(b.v[4] * (kGm1 / 2)) {}
  MCFD_HD double P(double pa) const { return (kGp1 / 2) * pa + p_const; }
  MCFD_HD double F(double pa) const {
    if (pa >= p) return (pa - p) / sqrt(rho * P(pa));
    return (pow(pa / p, (kGm1 / 2) / kGamma) - 1) * ((2 / kGm1) * a);
  }
  MCFD_HD double Fp(double pa) const {
    if (pa >= p) {
      double pp = P(pa);
      return (1 + (kGp1 / 4) * (p - pa) / pp) / sqrt(rho * pp);
    }
    return pow(pa / p, -(kGp1 / 2) / kGamma) / (rho * a);
  }
};
MCFD_HD State InsideFan(double gri_1, double gri_2) {  // exact.hpp:255-266
  constexpr double r = kGm1 / kGp1;
  double a = r * gri_2, a2 = a * a;
  double rho = pow(a2 / gri_1 * (1 / kGamma), 1 / kGm1);
  State s;
  s.v[0] = rho; s.v[1] = a; s.v[2] = 0; s.v[3] = 0; s.v[4] = a2 * rho * (1 / kGamma);
  return s;
}
// Toward the t-axis through wave 1 (left, sign = +1) or 3 (right, sign = -1).
MCFD_HD State NearShock(State const &before, State after, bool left) {  // exact.hpp:202-234
  double su = before.v[1];
  double divisor = (after.v[1] - before.v[1]) * before.v[0];
  su += (divisor != 0 ? (after.v[4] - before.v[4]) / divisor : before.v[1]);
  bool after_shock = left ? (su < 0) : (su > 0);
  if (after_shock) {
    after.v[0] = before.v[0];
    after.v[0] *= (before.v[1] - su) / (after.v[1] - su);
    return after;
  }
  return before;
}
MCFD_HD State NearFan(State const &before, State after, bool left) {  // exact.hpp:236-292
  double a_before = SoundSpeed(before.v[0], before.v[4]);
  double gri_1 = before.v[4] / pow(before.v[0], kGamma);
  double d = (2 / kGm1) * a_before;
  double gri_2 = left ? before.v[1] + d : before.v[1] - d;
  double e = (kGm1 / 2) * (before.v[1] - after.v[1]);
  double a_after = left ? a_before + e : a_before - e;
  bool axis_after = left ? (after.v[1] - a_after < 0) : (after.v[1] + a_after > 0);
  bool axis_before = left ? (before.v[1] - a_before > 0) : (before.v[1] + a_before < 0);
  if (axis_after) {
    after.v[0] = kGamma * after.v[4] / (a_after * a_after);
    return after;
  }
  if (axis_before) return before;
  return InsideFan(gri_1, gri_2);
}
MCFD_HD State NearVacuum(State const &l, State const &r) {  // exact.hpp:293-320
  double la = SoundSpeed(l.v[0], l.v[4]);
  if (l.v[1] > la) return l;
  double ra = SoundSpeed(r.v[0], r.v[4]);
  if (r.v[1] + ra < 0) return r;
  double gri_2 = l.v[1] + la * (2 / kGm1);
  if (gri_2 >= 0) return InsideFan(l.v[4] / pow(l.v[0], kGamma), gri_2);
  gri_2 = r.v[1] - ra * (2 / kGm1);
  if (gri_2 < 0) return InsideFan(r.v[4] / pow(r.v[0], kGamma), gri_2);
  State z;
  for (int i = 0; i < 5; ++i) z.v[i] = 0;
  return z;
}
// exact.hpp:36-43,62-138.  Every face point has its own solver object in
// the reference (fem.hpp:95-115), and that object keeps `star_u`, the star
// velocity of its last ordinary solve (exact.hpp:29,116), across calls.  The
// passive tangential velocities take the side given by its sign
// (exact.hpp:49-57), also in the vacuum branches, which do not update it.
// `star_positive` carries that sign in and out; a fresh solver holds 0.
MCFD_HD State ExactFlux(State const &l, State const &r, bool &star_positive) {
  bool sp = star_positive;
  State res;
  for (int i = 0; i < 5; ++i) res.v[i] = 0;
  bool done = false;
  if (l.v[0] > 0) {
    if (!(r.v[0] > 0)) {
      double al = SoundSpeed(l.v[0], l.v[4]);
      if (l.v[1] >= al) {
        res = l;
      } else {
        double us = l.v[1] + al * (2 / kGm1);
        if (us > 0) res = InsideFan(l.v[4] / pow(l.v[0], kGamma), us);
      }
      done = true;
    }
  } else {
    if (r.v[0] > 0) {
      double ar = SoundSpeed(r.v[0], r.v[4]);
      if (r.v[1] + ar <= 0) {
        res = r;
      } else {
        double us = r.v[1] - ar * (2 / kGm1);
        if (us < 0) res = InsideFan(r.v[4] / pow(r.v[0], kGamma), us);
      }
    }
    done = true;
  }
  if (!done) {
    ExactWave wl(l), wr(r);
    double du = r.v[1] - l.v[1];
    if (wl.F(0) + wr.F(0) + du < 0) {
      // exact.hpp:140-152 FindRoot: halve, then Newton.  The iteration caps
      // only bound pathological inputs (f(0) < 0 makes the halving stop
      // before x underflows; Newton converges in a few steps).
      double x = l.v[4];
      for (int it = 0; it < 2100 && wl.F(x) + wr.F(x) + du > 0; ++it) x *= 0.5;
      for (int it = 0; it < 200; ++it) {
        double f = wl.F(x) + wr.F(x) + du;
        if (!(f < -1e-8)) break;
        x -= f / (wl.Fp(x) + wr.Fp(x));
      }
      State star;
      star.v[0] = 0; star.v[2] = 0; star.v[3] = 0; star.v[4] = x;
      star.v[1] = 0.5 * (r.v[1] + wr.F(x) + l.v[1] - wl.F(x));
      sp = star.v[1] > 0;
      if (0 < star.v[1]) {
        res = (x >= l.v[4]) ? NearShock(l, star, true) : NearFan(l, star, true);
      } else {
        res = (x >= r.v[4]) ? NearShock(r, star, false) : NearFan(r, star, false);
      }
    } else {
      res = NearVacuum(l, r);
    }
  }
  res.v[2] = sp ? l.v[2] : r.v[2];
  res.v[3] = sp ? l.v[3] : r.v[3];
  star_positive = sp;
  return NormalFlux(res);
}

enum RiemannKind { kExact = 0, kHllc = 1, kRoe = 2 };

// rotated/euler.hpp:153-159: global conservative pair -> global flux.
// `star_positive` is the exact solver's per-point state (see ExactFlux); the
// other solvers keep none.
template <int kRiemann>
MCFD_HD State UpwindFlux(FaceFrame const &f, State const &cl, State const &cr, bool &star_positive) {
  State l = ToPrimitive(cl), r = ToPrimitive(cr);
  GlobalToNormal(f, &l);
  GlobalToNormal(f, &r);
  State flux;
  if constexpr (kRiemann == kExact) {
    flux = ExactFlux(l, r, star_positive);
  } else if constexpr (kRiemann == kHllc) {
    flux = HllcFlux(l, r);
  } else {
    flux = RoeFlux(l, r);
  }
  NormalToGlobal(f, &flux);
  return flux;
}

// rotated/euler.hpp:160-174 (default branch): (0, p n, 0).
MCFD_HD State WallFlux(FaceFrame const &f, State const &c) {
  State p = ToPrimitive(c);
  State flux;
  flux.v[0] = 0; flux.v[1] = p.v[4]; flux.v[2] = 0; flux.v[3] = 0; flux.v[4] = 0;
  NormalToGlobal(f, &flux);
  return flux;
}

// rotated/euler.hpp:144-151: physical normal flux of one state.
MCFD_HD State StateFlux(FaceFrame const &f, State const &c) {
  State p = ToPrimitive(c);
  GlobalToNormal(f, &p);
  State flux = NormalFlux(p);
  NormalToGlobal(f, &flux);
  return flux;
}

constexpr double kGasR = 287.05;  // riemann/euler/types.hpp:206-208

// rotated/euler.hpp:181-246 (Carlson 2011, 2.7 subsonic inflow), default
// branch: the boundary state from the interior's outgoing Riemann invariant
// and the given total conditions, given = (p_total, u_cos, v_cos, w_cos,
// T_total) with the flow direction cosines in global coordinates.
MCFD_HD State SubsonicInletFlux(FaceFrame const &f, State const &c, State const &given) {
  State p = ToPrimitive(c);
  const double c_in = SoundSpeed(p.v[0], p.v[4]);
  GlobalToNormal(f, &p);
  const double riemann_in = -p.v[1] - c_in * (2 / kGm1);
  const double u_cos = given.v[1], v_cos = given.v[2], w_cos = given.v[3], t_total = given.v[4];
  const double a = f.n[0] * u_cos + f.n[1] * v_cos + f.n[2] * w_cos;
  const double a2 = a * a;
  const double gamma_r = kGamma * kGasR;
  // larger root of qa x^2 + 2 qb x + qc = 0
  const double qa = a2 + 2 / kGm1, qb = riemann_in;
  const double qc = (kGm1 / 2) * (riemann_in * riemann_in) - a2 * gamma_r * t_total;
  const double sqrt_delta = sqrt(qb * qb - qa * qc);
  const double x1 = (-qb + sqrt_delta) / qa, x2 = (-qb - sqrt_delta) / qa;
  const double c_b = x1 < x2 ? x2 : x1;
  const double t_b = c_b * c_b / gamma_r;
  const double mach = sqrt((t_total / t_b - 1) / (kGm1 / 2));
  const double p_b = given.v[0] / pow(1 + (kGm1 / 2) * mach * mach, kGamma / kGm1);
  const double speed = mach * c_b;
  State b;
  b.v[4] = p_b;
  b.v[0] = p_b / kGasR / t_b;
  b.v[1] = u_cos * speed;
  b.v[2] = v_cos * speed;
  b.v[3] = w_cos * speed;
  GlobalToNormal(f, &b);
  State flux = NormalFlux(b);
  NormalToGlobal(f, &flux);
  return flux;
}

// rotated/euler.hpp:263-280 (Carlson 2011, 2.4 pressure outflow), default
// branch: the interior state at the given back pressure given[4], with the
// interior's p / rho kept.
MCFD_HD State SubsonicOutletFlux(FaceFrame const &f, State const &c, State const &given) {
  State p = ToPrimitive(c);
  const double rt = p.v[4] / p.v[0];
  p.v[4] = given.v[4];
  p.v[0] = p.v[4] / rt;
  GlobalToNormal(f, &p);
  State flux = NormalFlux(p);
  NormalToGlobal(f, &flux);
  return flux;
}

// Boundary kinds of the BC table (spatial/fem.hpp:196-222).
enum BcKind { kBcSmart = 0, kBcInviscidWall = 1, kBcSupersonicOutlet = 2, kBcSupersonicInlet = 3,
              kBcSubsonicInlet = 4, kBcSubsonicOutlet = 5 };

}  // namespace mcfd

#endif  // MINI_RIEMANN_EULER_CORE_HPP_
