// ec_state.hpp -- FP64 state algebra, reconstruction and Riemann fluxes of the
// euler_compare explicit update, for CDNA4 device code (and host code that
// needs the same arithmetic, e.g. initial data).
//
// Every formula keeps the reference's floating-point operation order so that
// a build without FMA contraction (-ffp-contract=off) reproduces the
// reference arithmetic operation by operation; citations are relative to
// amrex/apps/euler_compare/src/euler_compare/.  min/max follow std::min /
// std::max (AMReX_Algorithm.H:24-47), including their NaN behaviour, so they
// are written as ternaries, never fmin/fmax.
//
// The axis of a face is a run-time value here: x- and y-faces then share one
// instruction stream, so a wavefront that holds both kinds of face does not
// diverge.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace ec {

#define EC_HD __host__ __device__ __forceinline__

// Branch hints for the code layout: the fallback and boundary paths are
// placed out of line, so that the instruction fetch of the ordinary path
// stays sequential.  They change no result.
#define EC_LIKELY(x) __builtin_expect(!!(x), 1)
#define EC_UNLIKELY(x) __builtin_expect(!!(x), 0)

struct Prim { double r, u, v, p; };
struct Cons { double r, mx, my, E; };

EC_HD double smin(double a, double b) { return (b < a) ? b : a; }
EC_HD double smax(double a, double b) { return (a < b) ? b : a; }
EC_HD double clamp01(double x) { return smin(smax(x, 0.0), 1.0); }

// state_primitives.hpp:103-117
EC_HD Prim to_prim(const Cons& u, double gamma)
{
  const double inv_rho = 1.0 / u.r;
  const double ux = u.mx * inv_rho;
  const double uy = u.my * inv_rho;
  const double kinetic = 0.5 * u.r * (ux * ux + uy * uy);
  return Prim{u.r, ux, uy, (gamma - 1.0) * (u.E - kinetic)};
}

EC_HD Cons to_cons(const Prim& q, double gamma)
{
  const double kinetic = 0.5 * q.r * (q.u * q.u + q.v * q.v);
  return Cons{q.r, q.r * q.u, q.r * q.v, q.p / (gamma - 1.0) + kinetic};
}

// state_primitives.hpp:57-92
EC_HD double minmod(double a, double b)
{
  if (a * b <= 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

EC_HD double limited_slope(double backward, double forward, int limiter)
{
  if (backward * forward <= 0.0) return 0.0;
  if (limiter == 1) {
    return minmod(minmod(2.0 * backward, 0.5 * (backward + forward)), 2.0 * forward);
  }
  if (limiter == 2) {
    const double d = backward + forward;
    if (fabs(d) <= 1.0e-14) return 0.0;
    return 2.0 * backward * forward / d;
  }
  return minmod(backward, forward);
}

EC_HD double sound_speed(const Prim& q, double gamma) { return sqrt(gamma * q.p / q.r); }

EC_HD bool finite4(const Cons& f)
{
  return isfinite(f.r) && isfinite(f.mx) && isfinite(f.my) && isfinite(f.E);
}

EC_HD double normal_vel(const Prim& q, int axis) { return axis == 0 ? q.u : q.v; }

// reconstruction_fluxes.hpp:66-74
EC_HD Cons physical_flux(const Prim& q, int axis, double gamma)
{
  const Cons u = to_cons(q, gamma);
  if (axis == 0) return Cons{u.mx, u.mx * q.u + q.p, u.my * q.u, (u.E + q.p) * q.u};
  return Cons{u.my, u.mx * q.v, u.my * q.v + q.p, (u.E + q.p) * q.v};
}

// reconstruction_fluxes.hpp:76-92
EC_HD Cons rusanov(const Prim& L, const Prim& R, int axis, double gamma)
{
  const Cons ul = to_cons(L, gamma), ur = to_cons(R, gamma);
  const Cons fl = physical_flux(L, axis, gamma), fr = physical_flux(R, axis, gamma);
  const double s = smax(fabs(normal_vel(L, axis)) + sound_speed(L, gamma),
                        fabs(normal_vel(R, axis)) + sound_speed(R, gamma));
  const double hs = 0.5 * s;
  return Cons{0.5 * (fl.r + fr.r) - hs * (ur.r - ul.r), 0.5 * (fl.mx + fr.mx) - hs * (ur.mx - ul.mx),
              0.5 * (fl.my + fr.my) - hs * (ur.my - ul.my), 0.5 * (fl.E + fr.E) - hs * (ur.E - ul.E)};
}

// reconstruction_fluxes.hpp:117-149
EC_HD bool hllc_degenerate(const Prim& L, const Prim& R, int axis, double gamma)
{
  const double un_l = normal_vel(L, axis), un_r = normal_vel(R, axis);
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double s_l = smin(un_l - c_l, un_r - c_r);
  const double s_r = smax(un_l + c_l, un_r + c_r);
  if (!(s_l < s_r)) return true;
  if (s_l >= 0.0 || s_r <= 0.0) return false;
  const double den = L.r * (s_l - un_l) - R.r * (s_r - un_r);
  if (fabs(den) <= 1.0e-12) return true;
  const double s_m = (R.p - L.p + L.r * un_l * (s_l - un_l) - R.r * un_r * (s_r - un_r)) / den;
  return fabs(s_l - s_m) <= 1.0e-12 || fabs(s_r - s_m) <= 1.0e-12;
}

// reconstruction_fluxes.hpp:151-210 (star states :106-115).  Every Rusanov
// fallback of the reference is routed to one call site.
EC_HD Cons hllc(const Prim& L, const Prim& R, int axis, double gamma)
{
  const Cons ul = to_cons(L, gamma), ur = to_cons(R, gamma);
  const double un_l = normal_vel(L, axis), un_r = normal_vel(R, axis);
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double s_l = smin(un_l - c_l, un_r - c_r);
  const double s_r = smax(un_l + c_l, un_r + c_r);
  bool fallback = !(s_l < s_r);
  if (!fallback) {
    if (s_l >= 0.0) return physical_flux(L, axis, gamma);
    if (s_r <= 0.0) return physical_flux(R, axis, gamma);
    const double den = L.r * (s_l - un_l) - R.r * (s_r - un_r);
    fallback = fabs(den) <= 1.0e-12;
    if (!fallback) {
      const double s_m =
          (R.p - L.p + L.r * un_l * (s_l - un_l) - R.r * un_r * (s_r - un_r)) / den;
      const double lsd = s_l - s_m;
      const double rsd = s_r - s_m;
      fallback = fabs(lsd) <= 1.0e-12 || fabs(rsd) <= 1.0e-12;
      if (!fallback) {
        const double p_star_l = L.p + L.r * (s_l - un_l) * (s_m - un_l);
        const double p_star_r = R.p + R.r * (s_r - un_r) * (s_m - un_r);
        const double p_star = 0.5 * (p_star_l + p_star_r);
        // Only the upwind star state enters the flux (the reference builds
        // both and selects on s_m; the other has no effect on the result).
        const bool left = s_m >= 0.0;
        const Prim& K = left ? L : R;
        const Cons& uk = left ? ul : ur;
        const Cons fk = physical_flux(K, axis, gamma);
        const double sk = left ? s_l : s_r;
        const double unk = left ? un_l : un_r;
        const double kd = left ? lsd : rsd;
        const double rho_s = K.r * (sk - unk) / kd;
        const double e_s = ((sk - unk) * uk.E - K.p * unk + p_star * s_m) / kd;
        const double tang = axis == 0 ? K.v : K.u;
        const double smx = axis == 0 ? rho_s * s_m : rho_s * tang;
        const double smy = axis == 0 ? rho_s * tang : rho_s * s_m;
        return Cons{fk.r + sk * (rho_s - uk.r), fk.mx + sk * (smx - uk.mx),
                    fk.my + sk * (smy - uk.my), fk.E + sk * (e_s - uk.E)};
      }
    }
  }
  return rusanov(L, R, axis, gamma);
}

// reconstruction_fluxes.hpp:212-232
EC_HD void low_mach_correct(const Prim& L, const Prim& R, double gamma, Prim& cl, Prim& cr)
{
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double sp_l = sqrt(L.u * L.u + L.v * L.v);
  const double sp_r = sqrt(R.u * R.u + R.v * R.v);
  const double m_l = sp_l / smax(c_l, 1.0e-12);
  const double m_r = sp_r / smax(c_r, 1.0e-12);
  const double z = smin(smax(smax(m_l, m_r), 0.0), 1.0);
  const double u_avg = 0.5 * (L.u + R.u);
  const double v_avg = 0.5 * (L.v + R.v);
  cl = L;
  cr = R;
  cl.u = u_avg + 0.5 * z * (L.u - R.u);
  cr.u = u_avg + 0.5 * z * (R.u - L.u);
  cl.v = v_avg + 0.5 * z * (L.v - R.v);
  cr.v = v_avg + 0.5 * z * (R.v - L.v);
}

// reconstruction_fluxes.hpp:254-371 (Xie AM-HLLC-P).  Every HLLC fallback of
// the reference is routed to one call site.
EC_HD Cons xie_hllcp(const Prim& L, const Prim& R, int axis, double gamma, double sensor_f)
{
  const double q_l = normal_vel(L, axis), q_r = normal_vel(R, axis);
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double s_l = smin(q_l - c_l, q_r - c_r);
  const double s_r = smax(q_l + c_l, q_r + c_r);
  do {
    if (!(s_l < s_r)) break;
    if (s_l >= 0.0) return physical_flux(L, axis, gamma);
    if (s_r <= 0.0) return physical_flux(R, axis, gamma);
    const double alpha_l = L.r * (s_l - q_l);
    const double alpha_r = R.r * (s_r - q_r);
    const double alpha_denom = alpha_r - alpha_l;
    if (fabs(alpha_denom) <= 1.0e-12) break;
    const double s_star = (alpha_r * q_r - alpha_l * q_l + L.p - R.p) / alpha_denom;
    const double rld = s_l - s_star;
    const double rrd = s_r - s_star;
    if (fabs(rld) <= 1.0e-12 || fabs(rrd) <= 1.0e-12 || fabs(alpha_l) <= 1.0e-12 ||
        fabs(alpha_r) <= 1.0e-12) {
      break;
    }
    const double p_star =
        (alpha_r * L.p - alpha_l * R.p - alpha_l * alpha_r * (q_l - q_r)) / alpha_denom;
    const double sqrt_rho_l = sqrt(L.r);
    const double sqrt_rho_r = sqrt(R.r);
    const double roe_denom = sqrt_rho_l + sqrt_rho_r;
    if (!(roe_denom > 0.0)) break;
    const Cons ul = to_cons(L, gamma), ur = to_cons(R, gamma);
    const double u_hat = (sqrt_rho_l * L.u + sqrt_rho_r * R.u) / roe_denom;
    const double v_hat = (sqrt_rho_l * L.v + sqrt_rho_r * R.v) / roe_denom;
    const double q_hat = axis == 0 ? u_hat : v_hat;
    const double h_l = (ul.E + L.p) / L.r;
    const double h_r = (ur.E + R.p) / R.r;
    const double h_hat = (sqrt_rho_l * h_l + sqrt_rho_r * h_r) / roe_denom;
    const double c_hat_sq = (gamma - 1.0) * (h_hat - 0.5 * (u_hat * u_hat + v_hat * v_hat));
    if (!(c_hat_sq > 0.0) || !isfinite(c_hat_sq)) break;
    const double c_hat = sqrt(c_hat_sq);
    const double m_hat = q_hat / smax(c_hat, 1.0e-12);
    const double sp_l = sqrt(L.u * L.u + L.v * L.v);
    const double sp_r = sqrt(R.u * R.u + R.v * R.v);
    const double theta = clamp01(smax(sp_l / smax(c_l, 1.0e-12), sp_r / smax(c_r, 1.0e-12)));
    const double f = clamp01(sensor_f);
    const double pss = theta * p_star + (1.0 - theta) * 0.5 * (L.p + R.p);
    const double psss = f * pss + (1.0 - f) * p_star;
    const double beta_p = (f - 1.0) * s_l * s_r / (s_r - s_l) * (R.p - L.p) /
                          (c_hat_sq * (1.0 + fabs(m_hat)));
    // Upwind side selected on s_star (reconstruction_fluxes.hpp:362-365);
    // only its star density and energy are formed.
    const bool left = s_star >= 0.0;
    const Prim& K = left ? L : R;
    const double qn = left ? q_l : q_r;
    const double alpha = left ? alpha_l : alpha_r;
    const double rs = alpha / (left ? rld : rrd);
    const double ek = (left ? ul.E : ur.E) / K.r;
    const double es = ek + (s_star - qn) * (s_star + K.p / alpha);
    const double nx = axis == 0 ? 1.0 : 0.0;
    const double ny = axis == 1 ? 1.0 : 0.0;
    const Cons out{rs * s_star + beta_p,
                   rs * s_star * (K.u + nx * (s_star - qn)) + nx * psss + beta_p * u_hat,
                   rs * s_star * (K.v + ny * (s_star - qn)) + ny * psss + beta_p * v_hat,
                   s_star * (rs * es + p_star) + 0.5 * beta_p * (u_hat * u_hat + v_hat * v_hat)};
    if (!finite4(out)) break;
    return out;
  } while (false);
  return hllc(L, R, axis, gamma);
}

// reconstruction_fluxes.hpp:373-388; riemann is an ec_riemann value.
EC_HD Cons interface_flux(const Prim& L, const Prim& R, int axis, double gamma, int riemann,
                          double sensor_f)
{
  if (riemann == 0) return rusanov(L, R, axis, gamma);
  if (riemann == 2) {
    Prim cl, cr;
    low_mach_correct(L, R, gamma, cl, cr);
    return hllc(cl, cr, axis, gamma);
  }
  if (riemann == 3) return xie_hllcp(L, R, axis, gamma, sensor_f);
  return hllc(L, R, axis, gamma);
}

// explicit_update.hpp:21-30: the plain-HLLC star test, on the velocity-
// corrected states for the low-Mach HLLC.
EC_HD bool face_degenerate(const Prim& L, const Prim& R, int axis, double gamma, int riemann)
{
  if (riemann == 2) {
    Prim cl, cr;
    low_mach_correct(L, R, gamma, cl, cr);
    return hllc_degenerate(cl, cr, axis, gamma);
  }
  return hllc_degenerate(L, R, axis, gamma);
}

// reconstruction_fluxes.hpp:390-397.  min(a/b, b/a) == min(a,b)/max(a,b)
// bit for bit for finite positive a, b (the floor makes both so), which saves
// one division.
EC_HD double pressure_ratio(double pL, double pR)
{
  const double fl = 1.0e-14;
  const double a = isfinite(pL) ? smax(pL, fl) : fl;
  const double b = isfinite(pR) ? smax(pR, fl) : fl;
  return clamp01(smin(a, b) / smax(a, b));
}

}  // namespace ec
