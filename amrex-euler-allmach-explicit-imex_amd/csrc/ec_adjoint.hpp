// ec_adjoint.hpp -- hand-written reverse mode (vector-Jacobian products) of
// the explicit O2 update's pointwise functions.
//
// Every *_vjp recomputes the forward values it needs from its inputs, takes
// the branch the forward evaluation takes (wave selection, fallbacks, limiter
// cases, min/max arguments) and accumulates  xbar += (dy/dx)^T ybar  for the
// bar of its output.  Branch decisions are thereby frozen at the forward
// state, as an adjoint of a non-smooth scheme must be.  The forward formulas
// are those of ec_state.hpp (reconstruction_fluxes.hpp:66-371 and
// state_primitives.hpp:57-117 of the reference, relative to
// amrex/apps/euler_compare/src/euler_compare/).
#pragma once

#include "ec_state.hpp"

namespace ec {

EC_HD void add(Prim& a, const Prim& b)
{
  a.r += b.r; a.u += b.u; a.v += b.v; a.p += b.p;
}

EC_HD void add(Cons& a, const Cons& b)
{
  a.r += b.r; a.mx += b.mx; a.my += b.my; a.E += b.E;
}

// u = to_cons(q)
EC_HD void to_cons_vjp(const Prim& q, double gamma, const Cons& ub, Prim& qb)
{
  qb.r += ub.r + q.u * ub.mx + q.v * ub.my + 0.5 * (q.u * q.u + q.v * q.v) * ub.E;
  qb.u += q.r * ub.mx + q.r * q.u * ub.E;
  qb.v += q.r * ub.my + q.r * q.v * ub.E;
  qb.p += ub.E / (gamma - 1.0);
}

// q = to_prim(u)
EC_HD void to_prim_vjp(const Cons& u, double gamma, const Prim& qb, Cons& ub)
{
  const double inv = 1.0 / u.r;
  const double ux = u.mx * inv, uy = u.my * inv;
  const double Kb = -(gamma - 1.0) * qb.p;
  ub.E += (gamma - 1.0) * qb.p;
  double rb = qb.r + 0.5 * (ux * ux + uy * uy) * Kb;
  const double uxb = qb.u + u.r * ux * Kb;
  const double uyb = qb.v + u.r * uy * Kb;
  ub.mx += inv * uxb;
  ub.my += inv * uyb;
  const double invb = u.mx * uxb + u.my * uyb;
  rb += -inv * inv * invb;
  ub.r += rb;
}

// c = sqrt(gamma p / rho)
EC_HD void sound_speed_vjp(const Prim& q, double gamma, double cb, Prim& qb)
{
  const double c = sound_speed(q, gamma);
  if (c > 0.0) {
    qb.p += cb * gamma / (2.0 * q.r * c);
    qb.r += -cb * c / (2.0 * q.r);
  }
}

EC_HD void normal_vel_vjp(int axis, double b, Prim& qb)
{
  if (axis == 0) qb.u += b;
  else qb.v += b;
}

// f = physical_flux(q, axis)
EC_HD void physical_flux_vjp(const Prim& q, int axis, double gamma, const Cons& fb, Prim& qb)
{
  const Cons u = to_cons(q, gamma);
  Cons ub{0.0, 0.0, 0.0, 0.0};
  const double un = axis == 0 ? q.u : q.v;
  double unb = 0.0;
  if (axis == 0) {
    // (mx, mx u + p, my u, (E + p) u)
    ub.mx += fb.r + un * fb.mx;
    ub.my += un * fb.my;
    ub.E += un * fb.E;
    unb += u.mx * fb.mx + u.my * fb.my + (u.E + q.p) * fb.E;
    qb.p += fb.mx + un * fb.E;
  } else {
    // (my, mx v, my v + p, (E + p) v)
    ub.my += fb.r + un * fb.my;
    ub.mx += un * fb.mx;
    ub.E += un * fb.E;
    unb += u.mx * fb.mx + u.my * fb.my + (u.E + q.p) * fb.E;
    qb.p += fb.my + un * fb.E;
  }
  normal_vel_vjp(axis, unb, qb);
  to_cons_vjp(q, gamma, ub, qb);
}

// std::min(a, b) = (b < a) ? b : a ; std::max(a, b) = (a < b) ? b : a
EC_HD void min_vjp(double a, double b, double yb, double& ab, double& bb)
{
  if (b < a) bb += yb;
  else ab += yb;
}

EC_HD void max_vjp(double a, double b, double yb, double& ab, double& bb)
{
  if (a < b) bb += yb;
  else ab += yb;
}

EC_HD void rusanov_vjp(const Prim& L, const Prim& R, int axis, double gamma, const Cons& Fb,
                       Prim& Lb, Prim& Rb)
{
  const Cons ul = to_cons(L, gamma), ur = to_cons(R, gamma);
  const double unl = normal_vel(L, axis), unr = normal_vel(R, axis);
  const double cl = sound_speed(L, gamma), cr = sound_speed(R, gamma);
  const double al = fabs(unl) + cl, ar = fabs(unr) + cr;
  const double s = smax(al, ar);
  const Cons fhalf{0.5 * Fb.r, 0.5 * Fb.mx, 0.5 * Fb.my, 0.5 * Fb.E};
  physical_flux_vjp(L, axis, gamma, fhalf, Lb);
  physical_flux_vjp(R, axis, gamma, fhalf, Rb);
  const double hs = 0.5 * s;
  const double sb = -0.5 * (Fb.r * (ur.r - ul.r) + Fb.mx * (ur.mx - ul.mx) +
                            Fb.my * (ur.my - ul.my) + Fb.E * (ur.E - ul.E));
  to_cons_vjp(R, gamma, Cons{-hs * Fb.r, -hs * Fb.mx, -hs * Fb.my, -hs * Fb.E}, Rb);
  to_cons_vjp(L, gamma, Cons{hs * Fb.r, hs * Fb.mx, hs * Fb.my, hs * Fb.E}, Lb);
  double alb = 0.0, arb = 0.0;
  max_vjp(al, ar, sb, alb, arb);
  normal_vel_vjp(axis, (unl < 0.0 ? -1.0 : 1.0) * alb, Lb);
  normal_vel_vjp(axis, (unr < 0.0 ? -1.0 : 1.0) * arb, Rb);
  sound_speed_vjp(L, gamma, alb, Lb);
  sound_speed_vjp(R, gamma, arb, Rb);
}

EC_HD void hllc_vjp(const Prim& L, const Prim& R, int axis, double gamma, const Cons& Fb, Prim& Lb,
                    Prim& Rb)
{
  const Cons ul = to_cons(L, gamma), ur = to_cons(R, gamma);
  const double un_l = normal_vel(L, axis), un_r = normal_vel(R, axis);
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double a1 = un_l - c_l, a2 = un_r - c_r;
  const double b1 = un_l + c_l, b2 = un_r + c_r;
  const double s_l = smin(a1, a2);
  const double s_r = smax(b1, b2);
  if (!(s_l < s_r)) { rusanov_vjp(L, R, axis, gamma, Fb, Lb, Rb); return; }
  if (s_l >= 0.0) { physical_flux_vjp(L, axis, gamma, Fb, Lb); return; }
  if (s_r <= 0.0) { physical_flux_vjp(R, axis, gamma, Fb, Rb); return; }
  const double den = L.r * (s_l - un_l) - R.r * (s_r - un_r);
  if (fabs(den) <= 1.0e-12) { rusanov_vjp(L, R, axis, gamma, Fb, Lb, Rb); return; }
  const double num = R.p - L.p + L.r * un_l * (s_l - un_l) - R.r * un_r * (s_r - un_r);
  const double s_m = num / den;
  const double lsd = s_l - s_m, rsd = s_r - s_m;
  if (fabs(lsd) <= 1.0e-12 || fabs(rsd) <= 1.0e-12) { rusanov_vjp(L, R, axis, gamma, Fb, Lb, Rb); return; }
  const double psl = L.p + L.r * (s_l - un_l) * (s_m - un_l);
  const double psr = R.p + R.r * (s_r - un_r) * (s_m - un_r);
  const double ps = 0.5 * (psl + psr);
  const bool left = s_m >= 0.0;
  const Prim& K = left ? L : R;
  Prim& Kb = left ? Lb : Rb;
  const Cons& uk = left ? ul : ur;
  const double sk = left ? s_l : s_r;
  const double unk = left ? un_l : un_r;
  const double kd = left ? lsd : rsd;
  const double rho_s = K.r * (sk - unk) / kd;
  const double A = (sk - unk) * uk.E - K.p * unk + ps * s_m;
  const double e_s = A / kd;
  const double tang = axis == 0 ? K.v : K.u;
  const double smx = axis == 0 ? rho_s * s_m : rho_s * tang;
  const double smy = axis == 0 ? rho_s * tang : rho_s * s_m;
  // F = fK + sK (us - uK)
  physical_flux_vjp(K, axis, gamma, Fb, Kb);
  double skb = Fb.r * (rho_s - uk.r) + Fb.mx * (smx - uk.mx) + Fb.my * (smy - uk.my) +
               Fb.E * (e_s - uk.E);
  to_cons_vjp(K, gamma, Cons{-sk * Fb.r, -sk * Fb.mx, -sk * Fb.my, -sk * Fb.E}, Kb);
  const double smxb = sk * Fb.mx, smyb = sk * Fb.my;
  double rho_sb = sk * Fb.r, e_sb = sk * Fb.E, s_mb = 0.0, tangb = 0.0;
  if (axis == 0) {
    rho_sb += s_m * smxb + tang * smyb;
    s_mb += rho_s * smxb;
    tangb += rho_s * smyb;
  } else {
    rho_sb += tang * smxb + s_m * smyb;
    s_mb += rho_s * smyb;
    tangb += rho_s * smxb;
  }
  if (axis == 0) Kb.v += tangb;
  else Kb.u += tangb;
  // e_s = A / kd
  const double Ab = e_sb / kd;
  double kdb = -e_sb * e_s / kd;
  // A = (sk - unk) EK - pK unk + ps s_m
  double unkb = 0.0;
  skb += Ab * uk.E;
  unkb -= Ab * uk.E;
  to_cons_vjp(K, gamma, Cons{0.0, 0.0, 0.0, Ab * (sk - unk)}, Kb);
  Kb.p += -Ab * unk;
  unkb += -Ab * K.p;
  const double psb = Ab * s_m;
  s_mb += Ab * ps;
  // rho_s = K.r (sk - unk) / kd
  const double Bb = rho_sb / kd;
  kdb += -rho_sb * rho_s / kd;
  Kb.r += Bb * (sk - unk);
  skb += Bb * K.r;
  unkb -= Bb * K.r;
  // kd = sk - s_m
  skb += kdb;
  s_mb -= kdb;
  double s_lb = 0.0, s_rb = 0.0, un_lb = 0.0, un_rb = 0.0;
  if (left) { s_lb += skb; un_lb += unkb; } else { s_rb += skb; un_rb += unkb; }
  // ps = 0.5 (psl + psr)
  const double pslb = 0.5 * psb, psrb = 0.5 * psb;
  Lb.p += pslb;
  Lb.r += pslb * (s_l - un_l) * (s_m - un_l);
  s_lb += pslb * L.r * (s_m - un_l);
  s_mb += pslb * L.r * (s_l - un_l);
  un_lb += -pslb * L.r * (s_m - un_l) - pslb * L.r * (s_l - un_l);
  Rb.p += psrb;
  Rb.r += psrb * (s_r - un_r) * (s_m - un_r);
  s_rb += psrb * R.r * (s_m - un_r);
  s_mb += psrb * R.r * (s_r - un_r);
  un_rb += -psrb * R.r * (s_m - un_r) - psrb * R.r * (s_r - un_r);
  // s_m = num / den
  const double numb = s_mb / den;
  const double denb = -s_mb * s_m / den;
  Rb.p += numb;
  Lb.p -= numb;
  Lb.r += numb * un_l * (s_l - un_l);
  un_lb += numb * L.r * (s_l - un_l) - numb * L.r * un_l;
  s_lb += numb * L.r * un_l;
  Rb.r -= numb * un_r * (s_r - un_r);
  un_rb += -numb * R.r * (s_r - un_r) + numb * R.r * un_r;
  s_rb -= numb * R.r * un_r;
  // den = L.r (s_l - un_l) - R.r (s_r - un_r)
  Lb.r += denb * (s_l - un_l);
  s_lb += denb * L.r;
  un_lb -= denb * L.r;
  Rb.r -= denb * (s_r - un_r);
  s_rb -= denb * R.r;
  un_rb += denb * R.r;
  // s_l = min(un_l - c_l, un_r - c_r), s_r = max(un_l + c_l, un_r + c_r)
  double a1b = 0.0, a2b = 0.0, b1b = 0.0, b2b = 0.0;
  min_vjp(a1, a2, s_lb, a1b, a2b);
  max_vjp(b1, b2, s_rb, b1b, b2b);
  un_lb += a1b + b1b;
  un_rb += a2b + b2b;
  sound_speed_vjp(L, gamma, -a1b + b1b, Lb);
  sound_speed_vjp(R, gamma, -a2b + b2b, Rb);
  normal_vel_vjp(axis, un_lb, Lb);
  normal_vel_vjp(axis, un_rb, Rb);
}

// Corrected states of the low-Mach HLLC (reconstruction_fluxes.hpp:212-232).
EC_HD void low_mach_correct_vjp(const Prim& L, const Prim& R, double gamma, const Prim& clb,
                                const Prim& crb, Prim& Lb, Prim& Rb)
{
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double sp_l = sqrt(L.u * L.u + L.v * L.v), sp_r = sqrt(R.u * R.u + R.v * R.v);
  const double dl = smax(c_l, 1.0e-12), dr = smax(c_r, 1.0e-12);
  const double m_l = sp_l / dl, m_r = sp_r / dr;
  const double mm = smax(m_l, m_r);
  const double z0 = smax(mm, 0.0);
  const double z = smin(z0, 1.0);
  // rho and p pass through
  Lb.r += clb.r; Lb.p += clb.p; Rb.r += crb.r; Rb.p += crb.p;
  // cl.u = (L.u+R.u)/2 + z/2 (L.u - R.u), cr.u = (L.u+R.u)/2 + z/2 (R.u - L.u)
  const double hz = 0.5 * z;
  Lb.u += 0.5 * (clb.u + crb.u) + hz * (clb.u - crb.u);
  Rb.u += 0.5 * (clb.u + crb.u) - hz * (clb.u - crb.u);
  Lb.v += 0.5 * (clb.v + crb.v) + hz * (clb.v - crb.v);
  Rb.v += 0.5 * (clb.v + crb.v) - hz * (clb.v - crb.v);
  const double zb = 0.5 * ((L.u - R.u) * (clb.u - crb.u) + (L.v - R.v) * (clb.v - crb.v));
  // z = min(max(max(m_l, m_r), 0), 1)
  double z0b = 0.0, oneb = 0.0;
  min_vjp(z0, 1.0, zb, z0b, oneb);
  double mmb = 0.0, zerob = 0.0;
  max_vjp(mm, 0.0, z0b, mmb, zerob);
  double m_lb = 0.0, m_rb = 0.0;
  max_vjp(m_l, m_r, mmb, m_lb, m_rb);
  auto side = [&](const Prim& q, double sp, double c, double d, double mb, Prim& qb) {
    if (mb == 0.0) return;
    if (sp > 0.0) {
      qb.u += mb / d * q.u / sp;
      qb.v += mb / d * q.v / sp;
    }
    if (!(c < 1.0e-12)) sound_speed_vjp(q, gamma, -mb * sp / (d * d), qb);
  };
  side(L, sp_l, c_l, dl, m_lb, Lb);
  side(R, sp_r, c_r, dr, m_rb, Rb);
}

EC_HD double clamp01_vjp(double x, double yb)
{
  // clamp01(x) = min(max(x, 0), 1)
  const double a = smax(x, 0.0);
  double ab = 0.0, oneb = 0.0;
  min_vjp(a, 1.0, yb, ab, oneb);
  double xb = 0.0, zb = 0.0;
  max_vjp(x, 0.0, ab, xb, zb);
  return xb;
}

// Xie AM-HLLC-P (reconstruction_fluxes.hpp:254-371); sb receives the bar of
// the sensor value f.
EC_HD void xie_hllcp_vjp(const Prim& L, const Prim& R, int axis, double gamma, double sensor_f,
                         const Cons& Fb, Prim& Lb, Prim& Rb, double& sb)
{
  const double q_l = normal_vel(L, axis), q_r = normal_vel(R, axis);
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double a1 = q_l - c_l, a2 = q_r - c_r, b1 = q_l + c_l, b2 = q_r + c_r;
  const double s_l = smin(a1, a2);
  const double s_r = smax(b1, b2);
  bool fb = !(s_l < s_r);
  if (!fb) {
    if (s_l >= 0.0) { physical_flux_vjp(L, axis, gamma, Fb, Lb); return; }
    if (s_r <= 0.0) { physical_flux_vjp(R, axis, gamma, Fb, Rb); return; }
  }
  const double alpha_l = L.r * (s_l - q_l);
  const double alpha_r = R.r * (s_r - q_r);
  const double alpha_denom = alpha_r - alpha_l;
  fb = fb || fabs(alpha_denom) <= 1.0e-12;
  const double s_star = (alpha_r * q_r - alpha_l * q_l + L.p - R.p) / alpha_denom;
  const double rld = s_l - s_star, rrd = s_r - s_star;
  fb = fb || fabs(rld) <= 1.0e-12 || fabs(rrd) <= 1.0e-12 || fabs(alpha_l) <= 1.0e-12 ||
       fabs(alpha_r) <= 1.0e-12;
  const double p_star = (alpha_r * L.p - alpha_l * R.p - alpha_l * alpha_r * (q_l - q_r)) / alpha_denom;
  const double sql = sqrt(L.r), sqr = sqrt(R.r);
  const double roe = sql + sqr;
  fb = fb || !(roe > 0.0);
  const Cons ul = to_cons(L, gamma), ur = to_cons(R, gamma);
  const double u_hat = (sql * L.u + sqr * R.u) / roe;
  const double v_hat = (sql * L.v + sqr * R.v) / roe;
  const double q_hat = axis == 0 ? u_hat : v_hat;
  const double h_l = (ul.E + L.p) / L.r, h_r = (ur.E + R.p) / R.r;
  const double h_hat = (sql * h_l + sqr * h_r) / roe;
  const double c2 = (gamma - 1.0) * (h_hat - 0.5 * (u_hat * u_hat + v_hat * v_hat));
  fb = fb || !(c2 > 0.0) || !isfinite(c2);
  if (!fb) {
    const double c_hat = sqrt(c2);
    const double dch = smax(c_hat, 1.0e-12);
    const double m_hat = q_hat / dch;
    const double sp_l = sqrt(L.u * L.u + L.v * L.v), sp_r = sqrt(R.u * R.u + R.v * R.v);
    const double dl = smax(c_l, 1.0e-12), dr = smax(c_r, 1.0e-12);
    const double ml = sp_l / dl, mr = sp_r / dr;
    const double th_in = smax(ml, mr);
    const double theta = clamp01(th_in);
    const double f = clamp01(sensor_f);
    const double pav = 0.5 * (L.p + R.p);
    const double pss = theta * p_star + (1.0 - theta) * pav;
    const double psss = f * pss + (1.0 - f) * p_star;
    const double G1 = (f - 1.0) * s_l * s_r / (s_r - s_l);
    const double D = c2 * (1.0 + fabs(m_hat));
    const double beta_p = G1 * (R.p - L.p) / D;
    const bool left = s_star >= 0.0;
    const Prim& K = left ? L : R;
    const double qn = left ? q_l : q_r;
    const double alpha = left ? alpha_l : alpha_r;
    const double kd = left ? rld : rrd;
    const double rs = alpha / kd;
    const Cons& uk = left ? ul : ur;
    const double ek = uk.E / K.r;
    const double es = ek + (s_star - qn) * (s_star + K.p / alpha);
    const double nx = axis == 0 ? 1.0 : 0.0, ny = axis == 1 ? 1.0 : 0.0;
    const double uu = u_hat * u_hat + v_hat * v_hat;
    const Cons out{rs * s_star + beta_p,
                   rs * s_star * (K.u + nx * (s_star - qn)) + nx * psss + beta_p * u_hat,
                   rs * s_star * (K.v + ny * (s_star - qn)) + ny * psss + beta_p * v_hat,
                   s_star * (rs * es + p_star) + 0.5 * beta_p * uu};
    if (finite4(out)) {
      Prim& Kb = left ? Lb : Rb;
      const double tu = K.u + nx * (s_star - qn), tv = K.v + ny * (s_star - qn);
      // outputs
      const double rssb = Fb.r + tu * Fb.mx + tv * Fb.my;  // bar of rs*s_star
      double betab = Fb.r + u_hat * Fb.mx + v_hat * Fb.my + 0.5 * uu * Fb.E;
      double u_hb = beta_p * Fb.mx + beta_p * u_hat * Fb.E;
      double v_hb = beta_p * Fb.my + beta_p * v_hat * Fb.E;
      const double rsbtot = rs * s_star;
      double s_starb = rs * rssb + (rs * es + p_star) * Fb.E;
      double rsb = s_star * rssb + s_star * es * Fb.E;
      double esb = s_star * rs * Fb.E;
      double p_starb = s_star * Fb.E;
      double psssb = nx * Fb.mx + ny * Fb.my;
      // tu, tv
      Kb.u += rsbtot * Fb.mx;
      Kb.v += rsbtot * Fb.my;
      const double dsb = nx * rsbtot * Fb.mx + ny * rsbtot * Fb.my;  // bar of (s_star - qn)
      s_starb += dsb;
      double qnb = -dsb;
      // es = ek + (s_star - qn)(s_star + K.p/alpha)
      double ekb = esb;
      const double w2 = s_star + K.p / alpha;
      s_starb += esb * w2 + esb * (s_star - qn);
      qnb += -esb * w2;
      const double w2b = esb * (s_star - qn);
      Kb.p += w2b / alpha;
      double alphab = -w2b * K.p / (alpha * alpha);
      // ek = uk.E / K.r
      to_cons_vjp(K, gamma, Cons{0.0, 0.0, 0.0, ekb / K.r}, Kb);
      Kb.r += -ekb * ek / K.r;
      // rs = alpha / kd
      alphab += rsb / kd;
      const double kdb = -rsb * rs / kd;
      double s_lb = 0.0, s_rb = 0.0;
      if (left) s_lb += kdb; else s_rb += kdb;
      s_starb -= kdb;
      double alpha_lb = 0.0, alpha_rb = 0.0, q_lb = 0.0, q_rb = 0.0;
      if (left) { alpha_lb += alphab; q_lb += qnb; } else { alpha_rb += alphab; q_rb += qnb; }
      // beta_p = G1 (R.p - L.p) / D
      const double G1b = betab * (R.p - L.p) / D;
      Rb.p += betab * G1 / D;
      Lb.p -= betab * G1 / D;
      const double Db = -betab * beta_p / D;
      double c2b = Db * (1.0 + fabs(m_hat));
      const double m_hatb = Db * c2 * (m_hat < 0.0 ? -1.0 : 1.0);
      // G1 = (f-1) s_l s_r / (s_r - s_l)
      const double w = s_r - s_l;
      double fb_ = G1b * s_l * s_r / w;
      s_lb += G1b * (f - 1.0) * s_r / w + G1b * G1 / w;
      s_rb += G1b * (f - 1.0) * s_l / w - G1b * G1 / w;
      // psss = f pss + (1-f) p_star
      fb_ += psssb * (pss - p_star);
      const double pssb = psssb * f;
      p_starb += psssb * (1.0 - f);
      // pss = theta p_star + (1-theta) pav
      const double thetab = pssb * (p_star - pav);
      p_starb += pssb * theta;
      const double pavb = pssb * (1.0 - theta);
      Lb.p += 0.5 * pavb;
      Rb.p += 0.5 * pavb;
      sb += clamp01_vjp(sensor_f, fb_);
      // theta = clamp01(max(ml, mr))
      const double th_inb = clamp01_vjp(th_in, thetab);
      double mlb = 0.0, mrb = 0.0;
      max_vjp(ml, mr, th_inb, mlb, mrb);
      auto mach_side = [&](const Prim& q, double sp, double c, double d, double mb, Prim& qb) {
        if (mb == 0.0) return;
        if (sp > 0.0) {
          qb.u += mb / d * q.u / sp;
          qb.v += mb / d * q.v / sp;
        }
        if (!(c < 1.0e-12)) sound_speed_vjp(q, gamma, -mb * sp / (d * d), qb);
      };
      mach_side(L, sp_l, c_l, dl, mlb, Lb);
      mach_side(R, sp_r, c_r, dr, mrb, Rb);
      // m_hat = q_hat / max(c_hat, 1e-12)
      double q_hatb = m_hatb / dch;
      if (!(c_hat < 1.0e-12)) {
        const double c_hatb = -m_hatb * q_hat / (dch * dch);
        c2b += c_hatb / (2.0 * c_hat);
      }
      if (axis == 0) u_hb += q_hatb; else v_hb += q_hatb;
      // c2 = (g-1)(h_hat - uu/2)
      const double h_hatb = (gamma - 1.0) * c2b;
      u_hb += -(gamma - 1.0) * c2b * u_hat;
      v_hb += -(gamma - 1.0) * c2b * v_hat;
      // Roe averages x_hat = (sql xl + sqr xr)/roe
      double sqlb = 0.0, sqrb = 0.0;
      auto roe_avg = [&](double xl, double xr, double xh, double xhb, double& xlb, double& xrb) {
        xlb += xhb * sql / roe;
        xrb += xhb * sqr / roe;
        sqlb += xhb * (xl - xh) / roe;
        sqrb += xhb * (xr - xh) / roe;
      };
      double h_lb = 0.0, h_rb = 0.0;
      roe_avg(L.u, R.u, u_hat, u_hb, Lb.u, Rb.u);
      roe_avg(L.v, R.v, v_hat, v_hb, Lb.v, Rb.v);
      roe_avg(h_l, h_r, h_hat, h_hatb, h_lb, h_rb);
      // h = (E + p)/rho
      to_cons_vjp(L, gamma, Cons{0.0, 0.0, 0.0, h_lb / L.r}, Lb);
      Lb.p += h_lb / L.r;
      Lb.r += -h_lb * h_l / L.r;
      to_cons_vjp(R, gamma, Cons{0.0, 0.0, 0.0, h_rb / R.r}, Rb);
      Rb.p += h_rb / R.r;
      Rb.r += -h_rb * h_r / R.r;
      Lb.r += sqlb * 0.5 / sql;
      Rb.r += sqrb * 0.5 / sqr;
      // p_star = (alpha_r pl - alpha_l pr - alpha_l alpha_r (q_l - q_r)) / alpha_denom
      const double Pn = alpha_r * L.p - alpha_l * R.p - alpha_l * alpha_r * (q_l - q_r);
      const double Pnb = p_starb / alpha_denom;
      double adb = -p_starb * p_star / alpha_denom;
      alpha_rb += Pnb * (L.p - alpha_l * (q_l - q_r));
      alpha_lb += Pnb * (-R.p - alpha_r * (q_l - q_r));
      Lb.p += Pnb * alpha_r;
      Rb.p -= Pnb * alpha_l;
      q_lb -= Pnb * alpha_l * alpha_r;
      q_rb += Pnb * alpha_l * alpha_r;
      (void)Pn;
      // s_star = (alpha_r q_r - alpha_l q_l + pl - pr) / alpha_denom
      const double Snb = s_starb / alpha_denom;
      adb += -s_starb * s_star / alpha_denom;
      alpha_rb += Snb * q_r;
      q_rb += Snb * alpha_r;
      alpha_lb -= Snb * q_l;
      q_lb -= Snb * alpha_l;
      Lb.p += Snb;
      Rb.p -= Snb;
      // alpha_denom = alpha_r - alpha_l
      alpha_rb += adb;
      alpha_lb -= adb;
      // alpha_l = L.r (s_l - q_l), alpha_r = R.r (s_r - q_r)
      Lb.r += alpha_lb * (s_l - q_l);
      s_lb += alpha_lb * L.r;
      q_lb -= alpha_lb * L.r;
      Rb.r += alpha_rb * (s_r - q_r);
      s_rb += alpha_rb * R.r;
      q_rb -= alpha_rb * R.r;
      // wave speeds
      double a1b = 0.0, a2b = 0.0, b1b = 0.0, b2b = 0.0;
      min_vjp(a1, a2, s_lb, a1b, a2b);
      max_vjp(b1, b2, s_rb, b1b, b2b);
      q_lb += a1b + b1b;
      q_rb += a2b + b2b;
      sound_speed_vjp(L, gamma, -a1b + b1b, Lb);
      sound_speed_vjp(R, gamma, -a2b + b2b, Rb);
      normal_vel_vjp(axis, q_lb, Lb);
      normal_vel_vjp(axis, q_rb, Rb);
      return;
    }
  }
  hllc_vjp(L, R, axis, gamma, Fb, Lb, Rb);
}

EC_HD void interface_flux_vjp(const Prim& L, const Prim& R, int axis, double gamma, int riemann,
                              double sensor_f, const Cons& Fb, Prim& Lb, Prim& Rb, double& sb)
{
  if (riemann == 0) { rusanov_vjp(L, R, axis, gamma, Fb, Lb, Rb); return; }
  if (riemann == 2) {
    Prim cl, cr;
    low_mach_correct(L, R, gamma, cl, cr);
    Prim clb{0.0, 0.0, 0.0, 0.0}, crb{0.0, 0.0, 0.0, 0.0};
    hllc_vjp(cl, cr, axis, gamma, Fb, clb, crb);
    low_mach_correct_vjp(L, R, gamma, clb, crb, Lb, Rb);
    return;
  }
  if (riemann == 3) { xie_hllcp_vjp(L, R, axis, gamma, sensor_f, Fb, Lb, Rb, sb); return; }
  hllc_vjp(L, R, axis, gamma, Fb, Lb, Rb);
}

// slope = limited_slope(b, f): returns d slope / d b and d slope / d f.
EC_HD void limited_slope_grad(double b, double f, int limiter, double& db, double& df)
{
  db = df = 0.0;
  if (b * f <= 0.0) return;
  if (limiter == 1) {
    // minmod(minmod(2b, (b+f)/2), 2f)
    const double x = 2.0 * b, y = 0.5 * (b + f);
    double dx_ = 0.0, dy_ = 0.0;  // d inner / d x, d inner / d y
    double inner;
    if (x * y <= 0.0) { inner = 0.0; }
    else if (fabs(x) < fabs(y)) { inner = x; dx_ = 1.0; }
    else { inner = y; dy_ = 1.0; }
    const double z = 2.0 * f;
    if (inner * z <= 0.0) return;
    if (fabs(inner) < fabs(z)) {
      db = 2.0 * dx_ + 0.5 * dy_;
      df = 0.5 * dy_;
    } else {
      df = 2.0;
    }
    return;
  }
  if (limiter == 2) {
    const double d = b + f;
    if (fabs(d) <= 1.0e-14) return;
    const double s = 2.0 * b * f / d;
    db = (2.0 * f - s) / d;
    df = (2.0 * b - s) / d;
    return;
  }
  if (fabs(b) < fabs(f)) db = 1.0;
  else df = 1.0;
}

// ratio = pressure_ratio(pL, pR) = clamp01(min(a,b)/max(a,b)) with a, b the
// floored pressures: d ratio / d pL and d ratio / d pR along the forward's
// own argument selection (std::min picks a on ties, std::max picks a too, so
// a tie gives ratio a/a and a zero derivative).
EC_HD void pressure_ratio_grad(double pL, double pR, double& dl, double& dr)
{
  dl = dr = 0.0;
  const double fl = 1.0e-14;
  const double a = isfinite(pL) ? smax(pL, fl) : fl;
  const double b = isfinite(pR) ? smax(pR, fl) : fl;
  const bool lo_b = b < a, hi_b = a < b;
  const double lo = lo_b ? b : a, hi = hi_b ? b : a;
  const double x = lo / hi;
  if (!(x <= 1.0)) return;  // clamp01 saturated (cannot happen for lo <= hi)
  const double dlo = 1.0 / hi, dhi = -lo / (hi * hi);
  const double da = (lo_b ? 0.0 : dlo) + (hi_b ? 0.0 : dhi);
  const double db = (lo_b ? dlo : 0.0) + (hi_b ? dhi : 0.0);
  // a = max(pL, floor) passes dpL when pL is finite and not below the floor
  dl = (isfinite(pL) && !(pL < fl)) ? da : 0.0;
  dr = (isfinite(pR) && !(pR < fl)) ? db : 0.0;
}

}  // namespace ec
