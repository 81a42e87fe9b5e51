// ec_flux_window.hpp -- the four Riemann fluxes of ec_state.hpp, for face
// states inside the magnitude window of ec_exact_ops.hpp, with the same
// results bit for bit and fewer instructions:
//
//   * every division and square root goes through recip/quot/sqrt_core, and
//     one reciprocal serves each repeated denominator (rho_L and rho_R for
//     the sound speed, enthalpy and star energy; the HLLC-P alpha
//     denominator for s*, p* and the degenerate test's s_m; the Roe
//     denominator for u^, v^ and H^; gamma - 1 for both total energies;
//     the HLLC star denominator for rho* and E*);
//   * min / max / clamp01 are v_min_f64 / v_max_f64: inside the window no
//     operand is NaN and no min/max compares a -0 with a +0, so they equal the
//     reference's std::min / std::max ternaries;
//   * the non-finite checks the reference makes on c^2 and on the HLLC-P flux
//     are dropped: every intermediate is finite inside the window;
//   * the plain-HLLC degenerate-star test the reference's degenerate count
//     makes on the same face states (numerics/explicit_update.hpp:2-56,
//     reconstruction_fluxes.hpp:117-149) is evaluated inside the flux, from
//     its s_l, s_r and alpha values.
//
// The operation order of every reference formula is kept (citations are
// relative to amrex/apps/euler_compare/src/euler_compare/numerics/
// reconstruction_fluxes.hpp).  Fallback branches call the plain functions of
// ec_state.hpp, which take the reference's own path.
#pragma once

#include "ec_exact_ops.hpp"
#include "ec_state.hpp"

namespace ec {

// limited_slope with the minmod limiter (reconstruction_fluxes.hpp:2-12) for
// finite operands: when a * b > 0 both have one sign, and the operand of
// smaller magnitude (b on a tie, which then equals a) is that sign on
// min(|a|, |b|).  Three instructions fewer than the ternaries.
__device__ __forceinline__ double minmod_finite(double a, double b)
{
  const double m = copysign(fmin(fabs(a), fabs(b)), a);
  return a * b <= 0.0 ? 0.0 : m;
}

struct FaceSide {
  double E;   // total energy of the face state
  double q2;  // u^2 + v^2
  Recip rr;   // 1 / rho
  double c;   // sound speed
};

// to_conserved's energy and the sound speed of a face state
// (state_primitives.hpp:103-117, reconstruction_fluxes.hpp:62-64).
__device__ __forceinline__ FaceSide face_side(const Prim& q, double gamma, const Recip& rg)
{
  FaceSide s;
  s.q2 = q.u * q.u + q.v * q.v;
  s.E = quot_pos(q.p, rg) + 0.5 * q.r * s.q2;
  s.rr = recip(q.r);
  s.c = sqrt_core(quot_pos(gamma * q.p, s.rr));
  return s;
}

// physical_flux from the state's own energy (:66-74).
template <int AX>
__device__ __forceinline__ Cons phys_w(const Prim& q, double E)
{
  const double mx = q.r * q.u, my = q.r * q.v;
  if constexpr (AX == 0) return Cons{mx, mx * q.u + q.p, my * q.u, (E + q.p) * q.u};
  return Cons{my, mx * q.v, my * q.v + q.p, (E + q.p) * q.v};
}

// Rusanov (:76-92).
template <int AX>
__device__ __forceinline__ Cons rusanov_w(const Prim& L, const Prim& R, const FaceSide& a,
                                          const FaceSide& b)
{
  const Cons fl = phys_w<AX>(L, a.E), fr = phys_w<AX>(R, b.E);
  const double ql = AX == 0 ? L.u : L.v, qr = AX == 0 ? R.u : R.v;
  const double s = fmax(fabs(ql) + a.c, fabs(qr) + b.c);
  const double hs = 0.5 * s;
  const double mxl = L.r * L.u, myl = L.r * L.v, mxr = R.r * R.u, myr = R.r * R.v;
  return Cons{0.5 * (fl.r + fr.r) - hs * (R.r - L.r), 0.5 * (fl.mx + fr.mx) - hs * (mxr - mxl),
              0.5 * (fl.my + fr.my) - hs * (myr - myl), 0.5 * (fl.E + fr.E) - hs * (b.E - a.E)};
}

// HLLC (:151-210) with its degenerate-star test (:117-149), which equals the
// flux's own fallback test except that the supersonic cases count as
// non-degenerate.
template <int AX>
__device__ __forceinline__ Cons hllc_w(const Prim& L, const Prim& R, const FaceSide& a,
                                       const FaceSide& b, double gamma, bool& degen)
{
  const double un_l = AX == 0 ? L.u : L.v, un_r = AX == 0 ? R.u : R.v;
  const double s_l = fmin(un_l - a.c, un_r - b.c);
  const double s_r = fmax(un_l + a.c, un_r + b.c);
  bool fallback = !(s_l < s_r);
  degen = fallback;
  if (EC_LIKELY(!fallback)) {
    if (EC_UNLIKELY(s_l >= 0.0)) return phys_w<AX>(L, a.E);
    if (EC_UNLIKELY(s_r <= 0.0)) return phys_w<AX>(R, b.E);
    const double dl = s_l - un_l, dr = s_r - un_r;
    const double den = L.r * dl - R.r * dr;
    fallback = fabs(den) <= 1.0e-12;
    if (EC_LIKELY(!fallback)) {
      const double s_m = quot(R.p - L.p + L.r * un_l * dl - R.r * un_r * dr, recip(den));
      const double lsd = s_l - s_m;
      const double rsd = s_r - s_m;
      fallback = fabs(lsd) <= 1.0e-12 || fabs(rsd) <= 1.0e-12;
      if (EC_LIKELY(!fallback)) {
        const double p_star_l = L.p + L.r * dl * (s_m - un_l);
        const double p_star_r = R.p + R.r * dr * (s_m - un_r);
        const double p_star = 0.5 * (p_star_l + p_star_r);
        const bool left = s_m >= 0.0;
        const Prim K = left ? L : R;
        const double EK = left ? a.E : b.E;
        const double sk = left ? s_l : s_r;
        const double unk = left ? un_l : un_r;
        const double dk = left ? dl : dr;
        const Recip rk = recip(left ? lsd : rsd);
        const Cons fk = phys_w<AX>(K, EK);
        const double rho_s = quot(K.r * dk, rk);
        const double e_s = quot(dk * EK - K.p * unk + p_star * s_m, rk);
        const double tang = AX == 0 ? K.v : K.u;
        const double smx = AX == 0 ? rho_s * s_m : rho_s * tang;
        const double smy = AX == 0 ? rho_s * tang : rho_s * s_m;
        const double mxk = K.r * K.u, myk = K.r * K.v;
        return Cons{fk.r + sk * (rho_s - K.r), fk.mx + sk * (smx - mxk), fk.my + sk * (smy - myk),
                    fk.E + sk * (e_s - EK)};
      }
    }
    degen = true;
  }
  return rusanov_w<AX>(L, R, a, b);
}

// Low-Mach HLLC velocity blend (:212-232); the blended states keep rho and p,
// so their face quantities other than q2 carry over.
__device__ __forceinline__ void low_mach_w(const Prim& L, const Prim& R, const FaceSide& a,
                                           const FaceSide& b, Prim& cl, Prim& cr)
{
  const double sp_l = sqrt_core0(a.q2);
  const double sp_r = sqrt_core0(b.q2);
  const double m_l = quot_pos(sp_l, recip(fmax(a.c, 1.0e-12)));
  const double m_r = quot_pos(sp_r, recip(fmax(b.c, 1.0e-12)));
  const double z = fmin(fmax(fmax(m_l, m_r), 0.0), 1.0);
  const double u_avg = 0.5 * (L.u + R.u);
  const double v_avg = 0.5 * (L.v + R.v);
  cl = L;
  cr = R;
  cl.u = u_avg + 0.5 * z * (L.u - R.u);
  cr.u = u_avg + 0.5 * z * (R.u - L.u);
  cl.v = v_avg + 0.5 * z * (L.v - R.v);
  cr.v = v_avg + 0.5 * z * (R.v - L.v);
}

// Xie AM-HLLC-P (:254-371) with the plain-HLLC degenerate test on the same
// states.  The HLLC fallbacks run the reference's hllc_flux through
// ec_state.hpp.
template <int AX>
__device__ __forceinline__ Cons xie_hllcp_w(const Prim& L, const Prim& R, const FaceSide& a,
                                            const FaceSide& b, double gamma, double sensor_f,
                                            bool& degen)
{
  const double q_l = AX == 0 ? L.u : L.v, q_r = AX == 0 ? R.u : R.v;
  const double c_l = a.c, c_r = b.c;
  const double s_l = fmin(q_l - c_l, q_r - c_r);
  const double s_r = fmax(q_l + c_l, q_r + c_r);
  degen = true;
  do {
    if (EC_UNLIKELY(!(s_l < s_r))) break;
    degen = false;
    if (EC_UNLIKELY(s_l >= 0.0)) return phys_w<AX>(L, a.E);
    if (EC_UNLIKELY(s_r <= 0.0)) return phys_w<AX>(R, b.E);
    const double dl = s_l - q_l, dr = s_r - q_r;
    const double alpha_l = L.r * dl;
    const double alpha_r = R.r * dr;
    const double alpha_denom = alpha_r - alpha_l;
    degen = true;
    if (EC_UNLIKELY(fabs(alpha_denom) <= 1.0e-12)) break;
    const Recip ra = recip(alpha_denom);
    // hllc_star_state_is_degenerate: its denominator rho_l (s_l - u_l) -
    // rho_r (s_r - u_r) is -alpha_denom exactly, and a / -b == -(a / b).
    // (only |s_l - s_m| and |s_r - s_m| are used: a zero's sign cannot matter)
    const double s_m = -quot_pos(R.p - L.p + L.r * q_l * dl - R.r * q_r * dr, ra);
    degen = fabs(s_l - s_m) <= 1.0e-12 || fabs(s_r - s_m) <= 1.0e-12;
    const double s_star = quot(alpha_r * q_r - alpha_l * q_l + L.p - R.p, ra);
    const double rld = s_l - s_star;
    const double rrd = s_r - s_star;
    if (EC_UNLIKELY(fabs(rld) <= 1.0e-12 || fabs(rrd) <= 1.0e-12 || fabs(alpha_l) <= 1.0e-12 ||
                    fabs(alpha_r) <= 1.0e-12)) {
      break;
    }
    const double p_star = quot(alpha_r * L.p - alpha_l * R.p - alpha_l * alpha_r * (q_l - q_r), ra);
    const double w_l = sqrt_core(L.r);
    const double w_r = sqrt_core(R.r);
    const Recip rroe = recip(w_l + w_r);   // positive inside the window
    const double u_hat = quot(w_l * L.u + w_r * R.u, rroe);
    const double v_hat = quot(w_l * L.v + w_r * R.v, rroe);
    const double q_hat = AX == 0 ? u_hat : v_hat;
    const double h_l = quot_pos(a.E + L.p, a.rr);
    const double h_r = quot_pos(b.E + R.p, b.rr);
    const double h_hat = quot_pos(w_l * h_l + w_r * h_r, rroe);
    const double uv2 = u_hat * u_hat + v_hat * v_hat;
    const double c_hat_sq = (gamma - 1.0) * (h_hat - 0.5 * uv2);
    if (EC_UNLIKELY(!(c_hat_sq > 0.0))) break;         // finite inside the window
    const double c_hat = sqrt_core(c_hat_sq);
    const double m_hat = quot(q_hat, recip(fmax(c_hat, 1.0e-12)));
    const double sp_l = sqrt_core0(a.q2);
    const double sp_r = sqrt_core0(b.q2);
    const double theta = fmin(fmax(fmax(quot_pos(sp_l, recip(fmax(c_l, 1.0e-12))),
                                        quot_pos(sp_r, recip(fmax(c_r, 1.0e-12)))), 0.0), 1.0);
    const double f = fmin(fmax(sensor_f, 0.0), 1.0);
    const double pss = theta * p_star + (1.0 - theta) * 0.5 * (L.p + R.p);
    const double psss = f * pss + (1.0 - f) * p_star;
    const double beta_p = quot(quot((f - 1.0) * s_l * s_r, recip(s_r - s_l)) * (R.p - L.p),
                               recip(c_hat_sq * (1.0 + fabs(m_hat))));
    const bool left = s_star >= 0.0;
    const Prim K = left ? L : R;
    const double qn = left ? q_l : q_r;
    const double alpha = left ? alpha_l : alpha_r;
    const double EK = left ? a.E : b.E;
    const Recip rk = left ? a.rr : b.rr;
    const double rs = quot_pos(alpha, recip(left ? rld : rrd));   // |alpha| > 1e-12
    const double ek = quot_pos(EK, rk);
    const double es = ek + (s_star - qn) * (s_star + quot_pos(K.p, recip(alpha)));
    constexpr double nx = AX == 0 ? 1.0 : 0.0;
    constexpr double ny = AX == 1 ? 1.0 : 0.0;
    const double rss = rs * s_star;
    // finite inside the window, so the reference's non-finite fallback
    // cannot trigger
    return Cons{rss + beta_p, rss * (K.u + nx * (s_star - qn)) + nx * psss + beta_p * u_hat,
                rss * (K.v + ny * (s_star - qn)) + ny * psss + beta_p * v_hat,
                s_star * (rs * es + p_star) + 0.5 * beta_p * uv2};
  } while (false);
  return hllc(L, R, AX, gamma);
}

// interface_flux (:373-388) plus the face's degenerate-star flag
// (explicit_update.hpp:21-30; Rusanov counts none).
template <int AX, int RIEMANN>
__device__ __forceinline__ Cons interface_flux_w(const Prim& L, const Prim& R, double gamma,
                                                 const Recip& rg, double sensor_f, bool& degen)
{
  const FaceSide a = face_side(L, gamma, rg), b = face_side(R, gamma, rg);
  if constexpr (RIEMANN == 0) {
    degen = false;
    return rusanov_w<AX>(L, R, a, b);
  } else if constexpr (RIEMANN == 2) {
    Prim cl, cr;
    low_mach_w(L, R, a, b, cl, cr);
    FaceSide ca = a, cb = b;
    ca.q2 = cl.u * cl.u + cl.v * cl.v;
    cb.q2 = cr.u * cr.u + cr.v * cr.v;
    ca.E = quot_pos(cl.p, rg) + 0.5 * cl.r * ca.q2;
    cb.E = quot_pos(cr.p, rg) + 0.5 * cr.r * cb.q2;
    return hllc_w<AX>(cl, cr, ca, cb, gamma, degen);
  } else if constexpr (RIEMANN == 3) {
    return xie_hllcp_w<AX>(L, R, a, b, gamma, sensor_f, degen);
  } else {
    return hllc_w<AX>(L, R, a, b, gamma, degen);
  }
}

}  // namespace ec
