// ec_adjoint.hip -- adjoint of the explicit O2 step (device) and host-callable
// pointwise flux / VJP entry points used to verify the hand-written reverse
// mode against finite differences.
#include <cstring>
#include <limits>
#include <vector>

#include "ec_adjoint.hpp"
#include "ec_context.hpp"
#include "ec_ghost.hpp"

using namespace ec;

extern "C" ec_status ec_flux_host(const double* qL, const double* qR, int axis, double gamma,
                                  int riemann, double sensor_f, double* F)
{
  if (!qL || !qR || !F || axis < 0 || axis > 1 || riemann < 0 || riemann > 3) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  const Prim L{qL[0], qL[1], qL[2], qL[3]}, R{qR[0], qR[1], qR[2], qR[3]};
  const Cons f = interface_flux(L, R, axis, gamma, riemann, sensor_f);
  F[0] = f.r;
  F[1] = f.mx;
  F[2] = f.my;
  F[3] = f.E;
  return EC_OK;
}

extern "C" ec_status ec_flux_vjp_host(const double* qL, const double* qR, int axis, double gamma,
                                      int riemann, double sensor_f, const double* Fbar,
                                      double* qLbar, double* qRbar, double* sensor_bar)
{
  if (!qL || !qR || !Fbar || !qLbar || !qRbar || axis < 0 || axis > 1 || riemann < 0 ||
      riemann > 3) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  const Prim L{qL[0], qL[1], qL[2], qL[3]}, R{qR[0], qR[1], qR[2], qR[3]};
  Prim Lb{0.0, 0.0, 0.0, 0.0}, Rb{0.0, 0.0, 0.0, 0.0};
  double sb = 0.0;
  interface_flux_vjp(L, R, axis, gamma, riemann, sensor_f, Cons{Fbar[0], Fbar[1], Fbar[2], Fbar[3]},
                     Lb, Rb, sb);
  qLbar[0] = Lb.r; qLbar[1] = Lb.u; qLbar[2] = Lb.v; qLbar[3] = Lb.p;
  qRbar[0] = Rb.r; qRbar[1] = Rb.u; qRbar[2] = Rb.v; qRbar[3] = Rb.p;
  if (sensor_bar) *sensor_bar = sb;
  return EC_OK;
}

// ===================================================================== step
// Adjoint of one explicit step U' = U - dt div F(U) with dt and every branch
// frozen at U:  lambda = lambda' + (d(-dt div F)/dU)^T lambda'.
//   pass 1 (faces)   Fbar = dt/dh (lambda'_R - lambda'_L) (lambda' = 0 off the
//                    valid box), VJP through the flux into the bars of the two
//                    face states; for HLLC-P also the bar of the selected
//                    sensor ratio
//   pass 2 (ratios)  HLLC-P: each face ratio gathers its bar from the <= 5
//                    sensors that selected it and passes it to the pressure of
//                    its two face states
//   pass 3 (cells)   on the grown box [-2,n+1]^2: face-state bars -> slope bars
//                    -> primitive bars of the cell and its neighbours ->
//                    conserved bars (to_prim VJP)
//   pass 4 (fold)    each valid cell gathers the conserved bars of every
//                    grown cell whose ghost rule reads it (wrap, clamp, mirror
//                    with the momentum sign flipped; fixed states read none)
// Each pass writes every output exactly once: no atomics, deterministic.
namespace ec {

bool has_exact_side(const GridParams& g);

struct AdjWork {
  int nx = 0, ny = 0;
  double* states = nullptr;  // segment of forward states (compact), seg_cap of them
  int seg_cap = 0;
  double* lam = nullptr;     // two compact adjoint buffers
  double* dtlog = nullptr;   // device dt record
  int dt_cap = 0;
  double* qb = nullptr;    // 4 sides x 4 comps x GX*GY face-state bars
  double* ub = nullptr;    // 4 comps x GX*GY conserved bars of grown cells
  int* sel = nullptr;      // HLLC-P: sensor argmin per flux face
  double* mbar = nullptr;  // HLLC-P: bar of the selected ratio per flux face
  double* q = nullptr;     // 12 x GX*GY: primitives, x-slopes, y-slopes of grown cells
  // checkpointed sweep: segment recompute on its own stream, double-buffered
  hipStream_t rstream = nullptr;
  hipEvent_t ev_start = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_use[2] = {nullptr, nullptr};
};

void adjoint_release(ec_ctx* ctx)
{
  AdjWork* w = static_cast<AdjWork*>(ctx->adjoint);
  if (!w) return;
  if (w->qb) (void)hipFree(w->qb);
  if (w->ub) (void)hipFree(w->ub);
  if (w->sel) (void)hipFree(w->sel);
  if (w->mbar) (void)hipFree(w->mbar);
  if (w->q) (void)hipFree(w->q);
  if (w->states) (void)hipFree(w->states);
  if (w->lam) (void)hipFree(w->lam);
  if (w->dtlog) (void)hipFree(w->dtlog);
  for (hipEvent_t e : {w->ev_start, w->ev_rec[0], w->ev_rec[1], w->ev_use[0], w->ev_use[1]}) {
    if (e) (void)hipEventDestroy(e);
  }
  if (w->rstream) (void)hipStreamDestroy(w->rstream);
  delete w;
  ctx->adjoint = nullptr;
}

static ec_status adjoint_workspace(ec_ctx* ctx, AdjWork** out)
{
  AdjWork* w = static_cast<AdjWork*>(ctx->adjoint);
  if (!w) {
    w = new AdjWork();
    ctx->adjoint = w;
    w->nx = ctx->grid.nx;
    w->ny = ctx->grid.ny;
    const size_t G = (size_t)(w->nx + 4) * (w->ny + 4);
    const size_t nf = (size_t)(w->nx + 1) * w->ny + (size_t)w->nx * (w->ny + 1);
    bool ok = hipMalloc((void**)&w->qb, 16 * G * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->ub, 4 * G * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->sel, nf * sizeof(int)) == hipSuccess &&
              hipMalloc((void**)&w->mbar, nf * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&w->q, 12 * G * sizeof(double)) == hipSuccess;
    if (!ok) {
      adjoint_release(ctx);
      return EC_ERR_OUT_OF_MEMORY;
    }
  }
  *out = w;
  return EC_OK;
}

struct AdjArgs {
  const double* __restrict__ U;     // forward state (caller layout)
  const double* __restrict__ lam1;  // lambda' (caller layout)
  double* __restrict__ lam0;        // lambda (caller layout)
  long long rs, cs;
  GridParams g;
  double dt;
  double* qb;
  double* ub;
  int* sel;
  double* mbar;
  double* q;  // 12 x GX*GY: primitives, x-slopes, y-slopes (written by the prep pass)
};

// Face-state bar storage: side 0 E (x,+), 1 W (x,-), 2 N (y,+), 3 S (y,-).
EC_HD long long gidx(const GridParams& g, int i, int j)
{
  return (long long)(j + 2) * (g.nx + 4) + (i + 2);
}

EC_HD double* qb_at(const AdjArgs& A, int side, int i, int j)
{
  const long long G = (long long)(A.g.nx + 4) * (A.g.ny + 4);
  return A.qb + (long long)side * 4 * G + gidx(A.g, i, j);
}

// State of grown cell (i,j) through the explicit ghost rule, for a grid
// without exact-Dirichlet sides (the adjoint entry points reject those), so
// the exact solutions' code is not compiled into the adjoint kernels.
EC_HD Cons adj_load_cell(const AdjArgs& A, int i, int j)
{
  const GridParams& g = A.g;
  if (i < 0 && g.bc[0] == BC_FIXED) return fixed_state(g, 0);
  if (i >= g.nx && g.bc[1] == BC_FIXED) return fixed_state(g, 1);
  if (j < 0 && g.bc[2] == BC_FIXED) return fixed_state(g, 2);
  if (j >= g.ny && g.bc[3] == BC_FIXED) return fixed_state(g, 3);
  bool rx, ry;
  const int ii = map_axis(i, g.nx, g.bc[0], g.bc[1], rx);
  const int jj = map_axis(j, g.ny, g.bc[2], g.bc[3], ry);
  const long long k = (long long)jj * A.rs + ii;
  Cons c{A.U[k], A.U[A.cs + k], A.U[2 * A.cs + k], A.U[3 * A.cs + k]};
  if (rx) c.mx = -c.mx;
  if (ry) c.my = -c.my;
  return c;
}

// Primitive state of grown cell (i,j) through the explicit ghost rule.
EC_HD Prim prim_eval(const AdjArgs& A, int i, int j)
{
  return to_prim(adj_load_cell(A, i, j), A.g.gamma);
}

EC_HD Prim slope_eval(const AdjArgs& A, int i, int j, int axis, const Prim& c)
{
  const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
  const Prim m = prim_eval(A, i - di, j - dj), p = prim_eval(A, i + di, j + dj);
  const int lim = A.g.limiter;
  return Prim{limited_slope(c.r - m.r, p.r - c.r, lim), limited_slope(c.u - m.u, p.u - c.u, lim),
              limited_slope(c.v - m.v, p.v - c.v, lim), limited_slope(c.p - m.p, p.p - c.p, lim)};
}

EC_HD long long grown_count(const GridParams& g) { return (long long)(g.nx + 4) * (g.ny + 4); }

EC_HD Prim qslot(const AdjArgs& A, int block, int i, int j)
{
  const long long G = grown_count(A.g);
  const double* b = A.q + (long long)block * 4 * G + gidx(A.g, i, j);
  return Prim{b[0], b[G], b[2 * G], b[3 * G]};
}

// Prep pass: the primitives of every grown cell and, for O2, the limited
// slopes of the cells that own face states ([-1,nx] x [-1,ny]), computed once
// per step (the forward step's stages A and B) for the passes below.
EC_HD void adj_prep_item(const AdjArgs& A, long long t)
{
  const GridParams& g = A.g;
  const int GX = g.nx + 4;
  const long long G = grown_count(g);
  const int i = (int)(t % GX) - 2, j = (int)(t / GX) - 2;
  const Prim c = prim_eval(A, i, j);
  double* q = A.q + t;
  q[0] = c.r; q[G] = c.u; q[2 * G] = c.v; q[3 * G] = c.p;
  if (g.order == 2 && i >= -1 && j >= -1 && i <= g.nx && j <= g.ny) {
    for (int axis = 0; axis < 2; ++axis) {
      const Prim s = slope_eval(A, i, j, axis, c);
      double* d = A.q + (long long)(4 + 4 * axis) * G + t;
      d[0] = s.r; d[G] = s.u; d[2 * G] = s.v; d[3 * G] = s.p;
    }
  }
}

EC_HD Prim prim_at(const AdjArgs& A, int i, int j) { return qslot(A, 0, i, j); }

EC_HD Prim face_state_at(const AdjArgs& A, int i, int j, int axis, double hs)
{
  Prim q = prim_at(A, i, j);
  if (A.g.order == 2) {
    const Prim s = qslot(A, axis == 0 ? 1 : 2, i, j);
    q.r = q.r + hs * s.r;
    q.u = q.u + hs * s.u;
    q.v = q.v + hs * s.v;
    q.p = q.p + hs * s.p;
  }
  return q;
}

EC_HD Cons lam_at(const AdjArgs& A, int i, int j)
{
  if (i < 0 || j < 0 || i >= A.g.nx || j >= A.g.ny) return Cons{0.0, 0.0, 0.0, 0.0};
  const long long k = (long long)j * A.rs + i;
  return Cons{A.lam1[k], A.lam1[A.cs + k], A.lam1[2 * A.cs + k], A.lam1[3 * A.cs + k]};
}

EC_HD double ratio_of(const AdjArgs& A, int fi, int fj, int axis)
{
  const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
  return pressure_ratio(face_state_at(A, fi - di, fj - dj, axis, 0.5).p,
                        face_state_at(A, fi, fj, axis, -0.5).p);
}

// One flux face's VJP; the axis and the Riemann solver are compile-time
// constants (RIEM < 0: read from the grid at run time, the host path).
template <int axis, int RIEM>
EC_HD void adj_face_body(const AdjArgs& A, long long t, int fi, int fj)
{
  const GridParams& g = A.g;
  {
    const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
    const int li = fi - di, lj = fj - dj;
    const int riemann = RIEM >= 0 ? RIEM : g.riemann;
    const double s = (axis == 0 ? g.inv_dx : g.inv_dy) * A.dt;
    const Cons lr = lam_at(A, fi, fj), ll = lam_at(A, li, lj);
    const Cons Fb{s * (lr.r - ll.r), s * (lr.mx - ll.mx), s * (lr.my - ll.my), s * (lr.E - ll.E)};
    Prim Lb{0.0, 0.0, 0.0, 0.0}, Rb{0.0, 0.0, 0.0, 0.0};
    int sel = -1;
    double mb = 0.0;
    if (Fb.r != 0.0 || Fb.mx != 0.0 || Fb.my != 0.0 || Fb.E != 0.0) {
      const Prim L = face_state_at(A, li, lj, axis, 0.5);
      const Prim R = face_state_at(A, fi, fj, axis, -0.5);
      double sensor = 1.0, m = 1.0;
      if (riemann == 3) {
        // reconstruction_fluxes.hpp:421-440, the first minimal ratio wins
        double r[5];
        if (axis == 0) {
          r[0] = ratio_of(A, fi, fj, 0);
          r[1] = ratio_of(A, fi - 1, fj, 1);
          r[2] = ratio_of(A, fi - 1, fj + 1, 1);
          r[3] = ratio_of(A, fi, fj, 1);
          r[4] = ratio_of(A, fi, fj + 1, 1);
        } else {
          r[0] = ratio_of(A, fi, fj, 1);
          r[1] = ratio_of(A, fi, fj - 1, 0);
          r[2] = ratio_of(A, fi + 1, fj - 1, 0);
          r[3] = ratio_of(A, fi, fj, 0);
          r[4] = ratio_of(A, fi + 1, fj, 0);
        }
        m = r[0];
        sel = 0;
        for (int k = 1; k < 5; ++k) {
          if (r[k] < m) {
            m = r[k];
            sel = k;
          }
        }
        const double c = clamp01(m);
        sensor = c * c * c;
      }
      double sb = 0.0;
      interface_flux_vjp(L, R, axis, g.gamma, riemann, sensor, Fb, Lb, Rb, sb);
      if (riemann == 3) {
        const double c = clamp01(m);
        mb = clamp01_vjp(m, 3.0 * c * c * sb);
      }
    }
    double* qL = qb_at(A, axis == 0 ? 0 : 2, li, lj);
    double* qR = qb_at(A, axis == 0 ? 1 : 3, fi, fj);
    const long long G = (long long)(g.nx + 4) * (g.ny + 4);
    qL[0] = Lb.r; qL[G] = Lb.u; qL[2 * G] = Lb.v; qL[3 * G] = Lb.p;
    qR[0] = Rb.r; qR[G] = Rb.u; qR[2 * G] = Rb.v; qR[3 * G] = Rb.p;
    if (A.sel) {
      A.sel[t] = sel;
      A.mbar[t] = mb;
    }
  }
}

// Number of items of the face pass: every flux face, plus, for HLLC-P, the
// faces outside the flux set whose pressure ratios the 5-face sensor reads
// (x-faces of rows -1 and ny, y-faces of columns -1 and nx).
EC_HD long long face_items(const GridParams& g)
{
  const long long nf = (long long)(g.nx + 1) * g.ny + (long long)g.nx * (g.ny + 1);
  return g.riemann == 3 ? nf + 2LL * (g.nx + 1) + 2LL * (g.ny + 1) : nf;
}

EC_HD void zero_slot(const AdjArgs& A, int side, int i, int j)
{
  const long long G = grown_count(A.g);
  double* p = qb_at(A, side, i, j);
  p[0] = 0.0; p[G] = 0.0; p[2 * G] = 0.0; p[3 * G] = 0.0;
}

template <int RIEM = -1>
EC_HD void adj_face_item(const AdjArgs& A, long long t)
{
  const GridParams& g = A.g;
  const long long nxf = (long long)(g.nx + 1) * g.ny;
  const long long nf = nxf + (long long)g.nx * (g.ny + 1);
  if (t >= nf) {
    // a ratio-only face: its face-state bars start at zero (the ratio pass
    // adds the pressure part)
    const long long e = t - nf;
    if (e < 2LL * (g.nx + 1)) {
      const int fi = (int)(e % (g.nx + 1)), fj = e / (g.nx + 1) == 0 ? -1 : g.ny;
      zero_slot(A, 0, fi - 1, fj);
      zero_slot(A, 1, fi, fj);
    } else {
      const long long e2 = e - 2LL * (g.nx + 1);
      const int fj = (int)(e2 % (g.ny + 1)), fi = e2 / (g.ny + 1) == 0 ? -1 : g.nx;
      zero_slot(A, 2, fi, fj - 1);
      zero_slot(A, 3, fi, fj);
    }
    return;
  }
  // the axis is a compile-time constant of each body (its selects fold away)
  if (t < nxf) {
    adj_face_body<0, RIEM>(A, t, (int)(t % (g.nx + 1)), (int)(t / (g.nx + 1)));
  } else {
    adj_face_body<1, RIEM>(A, t, (int)((t - nxf) % g.nx), (int)((t - nxf) / g.nx));
  }
}

// The adjoint passes of a step are the sweep's critical path; the segment
// recompute on the second stream is not (it runs a segment ahead).  Their
// waves run at priority 3, above every stage of the explicit step kernel
// (0 and 1), so that a SIMD holding both issues the adjoint first
// (checkpointed sweep at 640x160, 768 steps: 33.2 -> 32.8 ms).
__device__ __forceinline__ void adj_priority() { __builtin_amdgcn_s_setprio(3); }

__global__ void __launch_bounds__(256) adj_prep_kernel(AdjArgs A)
{
  adj_priority();
  const long long G = grown_count(A.g);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < G;
       t += (long long)gridDim.x * blockDim.x) {
    adj_prep_item(A, t);
  }
}

template <int RIEM>
__global__ void __launch_bounds__(256) adj_faces_kernel(AdjArgs A)
{
  adj_priority();
  const long long nf = face_items(A.g);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nf;
       t += (long long)gridDim.x * blockDim.x) {
    adj_face_item<RIEM>(A, t);
  }
}

// HLLC-P: gather each face ratio's bar and pass it to its face pressures.
EC_HD void adj_ratio_item(const AdjArgs& A, long long t)
{
  const GridParams& g = A.g;
  const long long nxf = (long long)(g.nx + 1) * g.ny;
  const long long nxr = (long long)(g.nx + 1) * (g.ny + 2), nyr = (long long)(g.nx + 2) * (g.ny + 1);
  auto xf = [&](int fi, int j, int want) -> double {  // x flux face (fi, j)
    (void)nxr;
    if (fi < 0 || fi > g.nx || j < 0 || j >= g.ny) return 0.0;
    const long long t = (long long)j * (g.nx + 1) + fi;
    return A.sel[t] == want ? A.mbar[t] : 0.0;
  };
  auto yf = [&](int i, int fj, int want) -> double {  // y flux face (i, fj)
    if (i < 0 || i >= g.nx || fj < 0 || fj > g.ny) return 0.0;
    const long long t = nxf + (long long)fj * g.nx + i;
    return A.sel[t] == want ? A.mbar[t] : 0.0;
  };
  const long long G = (long long)(g.nx + 4) * (g.ny + 4);
  (void)nyr;
  {
    int axis, fi, fj;
    double bar;
    if (t < nxr) {
      axis = 0;
      fi = (int)(t % (g.nx + 1));
      fj = (int)(t / (g.nx + 1)) - 1;
      bar = xf(fi, fj, 0) + yf(fi, fj + 1, 1) + yf(fi - 1, fj + 1, 2) + yf(fi, fj, 3) +
            yf(fi - 1, fj, 4);
    } else {
      axis = 1;
      const long long s = t - nxr;
      fi = (int)(s % (g.nx + 2)) - 1;
      fj = (int)(s / (g.nx + 2));
      bar = yf(fi, fj, 0) + xf(fi + 1, fj, 1) + xf(fi + 1, fj - 1, 2) + xf(fi, fj, 3) +
            xf(fi, fj - 1, 4);
    }
    if (bar == 0.0) return;
    const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
    const double pl = face_state_at(A, fi - di, fj - dj, axis, 0.5).p;
    const double pr = face_state_at(A, fi, fj, axis, -0.5).p;
    double dl, dr;
    pressure_ratio_grad(pl, pr, dl, dr);
    qb_at(A, axis == 0 ? 0 : 2, fi - di, fj - dj)[3 * G] += bar * dl;
    qb_at(A, axis == 0 ? 1 : 3, fi, fj)[3 * G] += bar * dr;
  }
}

__global__ void __launch_bounds__(256) adj_ratios_kernel(AdjArgs A)
{
  adj_priority();
  const long long nr = (long long)(A.g.nx + 1) * (A.g.ny + 2) + (long long)(A.g.nx + 2) * (A.g.ny + 1);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nr;
       t += (long long)gridDim.x * blockDim.x) {
    adj_ratio_item(A, t);
  }
}

// Whether the face pass wrote slot (side, i, j) this step; every other slot
// reads as zero (no face state of that side exists there).
EC_HD bool slot_written(const GridParams& g, int side, int i, int j)
{
  const bool hp = g.riemann == 3;
  if (side < 2) {
    const bool row = (j >= 0 && j < g.ny) || (hp && (j == -1 || j == g.ny));
    return row && (side == 0 ? (i >= -1 && i < g.nx) : (i >= 0 && i <= g.nx));
  }
  const bool col = (i >= 0 && i < g.nx) || (hp && (i == -1 || i == g.nx));
  return col && (side == 2 ? (j >= -1 && j < g.ny) : (j >= 0 && j <= g.ny));
}

EC_HD Prim qb_load(const AdjArgs& A, int side, int i, int j)
{
  if (!slot_written(A.g, side, i, j)) return Prim{0.0, 0.0, 0.0, 0.0};
  const long long G = (long long)(A.g.nx + 4) * (A.g.ny + 4);
  const double* p = qb_at(A, side, i, j);
  return Prim{p[0], p[G], p[2 * G], p[3 * G]};
}

// Bar of the slope of cell (i,j) along an axis: 0.5 (bar_plus_state - bar_minus_state).
EC_HD Prim slope_bar(const AdjArgs& A, int i, int j, int axis)
{
  const Prim a = qb_load(A, axis == 0 ? 0 : 2, i, j), b = qb_load(A, axis == 0 ? 1 : 3, i, j);
  return Prim{0.5 * (a.r - b.r), 0.5 * (a.u - b.u), 0.5 * (a.v - b.v), 0.5 * (a.p - b.p)};
}

// Valid cells at least 2 cells from every side: no ghost cell is read from
// them, so their adjoint is complete after the cell pass.
EC_HD bool fold_interior(const GridParams& g, int i, int j)
{
  return i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3;
}

EC_HD void adj_cell_item(const AdjArgs& A, long long t)
{
  const GridParams& g = A.g;
  const int GX = g.nx + 4;
  const long long G = (long long)GX * (g.ny + 4);
  {
    const int i = (int)(t % GX) - 2, j = (int)(t / GX) - 2;
    Prim qbar{0.0, 0.0, 0.0, 0.0};
    for (int side = 0; side < 4; ++side) add(qbar, qb_load(A, side, i, j));
    if (g.order == 2) {
      for (int axis = 0; axis < 2; ++axis) {
        const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
        // slopes of (i,j) and of its two neighbours on this axis read q(i,j)
        for (int o = -1; o <= 1; ++o) {
          const int ci = i + o * di, cj = j + o * dj;
          if (ci < -1 || cj < -1 || ci > g.nx || cj > g.ny) continue;  // no face states there
          const Prim sb = slope_bar(A, ci, cj, axis);
          if (sb.r == 0.0 && sb.u == 0.0 && sb.v == 0.0 && sb.p == 0.0) continue;
          const Prim qm = prim_at(A, ci - di, cj - dj), qc = prim_at(A, ci, cj),
                     qp = prim_at(A, ci + di, cj + dj);
          const double bk[4] = {qc.r - qm.r, qc.u - qm.u, qc.v - qm.v, qc.p - qm.p};
          const double fw[4] = {qp.r - qc.r, qp.u - qc.u, qp.v - qc.v, qp.p - qc.p};
          const double sbv[4] = {sb.r, sb.u, sb.v, sb.p};
          double contrib[4];
          for (int c = 0; c < 4; ++c) {
            double db, df;
            limited_slope_grad(bk[c], fw[c], g.limiter, db, df);
            // slope(ci) = lim(q_ci - q_{ci-1}, q_{ci+1} - q_ci)
            double d = 0.0;
            if (o == 0) d = db - df;       // (i,j) is the centre
            else if (o == 1) d = -db;      // (i,j) = centre - 1
            else d = df;                   // (i,j) = centre + 1
            contrib[c] = sbv[c] * d;
          }
          qbar.r += contrib[0];
          qbar.u += contrib[1];
          qbar.v += contrib[2];
          qbar.p += contrib[3];
        }
      }
    }
    Cons ubar{0.0, 0.0, 0.0, 0.0};
    if (qbar.r != 0.0 || qbar.u != 0.0 || qbar.v != 0.0 || qbar.p != 0.0) {
      to_prim_vjp(adj_load_cell(A, i, j), g.gamma, qbar, ubar);
    }
    if (fold_interior(g, i, j)) {
      // no ghost cell reads (i,j): lambda_n = lambda_next + its own bar
      const long long k = (long long)j * A.rs + i;
      const double acc[4] = {0.0 + ubar.r, 0.0 + ubar.mx, 0.0 + ubar.my, 0.0 + ubar.E};
      for (int c = 0; c < 4; ++c) A.lam0[c * A.cs + k] = A.lam1[c * A.cs + k] + acc[c];
      return;
    }
    A.ub[t] = ubar.r;
    A.ub[G + t] = ubar.mx;
    A.ub[2 * G + t] = ubar.my;
    A.ub[3 * G + t] = ubar.E;
  }
}

// The cell pass as LDS tiles of CTX x CTY grown cells (one per thread): the
// primitives of the tile plus a 2-cell ring, and the slope bars and summed
// face-state bars of the tile plus a 1-cell ring, are staged once, so each
// limiter gradient and slot is read from LDS rather than recomputed per
// neighbour, and every face-state bar slot is read from memory once.  Same
// arithmetic, in the same order, as adj_cell_item.
constexpr int CTX = 32, CTY = 8, CNT = CTX * CTY;
__global__ void __launch_bounds__(CNT) adj_cells_tile_kernel(AdjArgs A)
{
  adj_priority();
  constexpr int QX = CTX + 4, QY = CTY + 4, NQ = QX * QY;  // cells [-2, CT+1]
  constexpr int BX = CTX + 2, BY = CTY + 2, NB = BX * BY;  // cells [-1, CT]
  __shared__ double qs[4][NQ];
  __shared__ double sbs[12][NB];  // x slope bars, y slope bars, summed face-state bars
  const GridParams& g = A.g;
  const int GX = g.nx + 4, GY = g.ny + 4;
  const int tiles_x = (GX + CTX - 1) / CTX;
  const int gi0 = (int)(blockIdx.x % tiles_x) * CTX - 2, gj0 = (int)(blockIdx.x / tiles_x) * CTY - 2;
  const long long G = grown_count(g);
  const bool o2 = g.order == 2;
  if (o2) {
    for (int t = threadIdx.x; t < NQ; t += CNT) {
      const int i = gi0 - 2 + t % QX, j = gj0 - 2 + t / QX;
      if (i >= -2 && j >= -2 && i <= g.nx + 1 && j <= g.ny + 1) {
        const long long k = gidx(g, i, j);
        qs[0][t] = A.q[k];
        qs[1][t] = A.q[G + k];
        qs[2][t] = A.q[2 * G + k];
        qs[3][t] = A.q[3 * G + k];
      }
    }
    for (int t = threadIdx.x; t < NB; t += CNT) {
      const int i = gi0 - 1 + t % BX, j = gj0 - 1 + t / BX;
      const Prim e = qb_load(A, 0, i, j), w = qb_load(A, 1, i, j);
      const Prim n = qb_load(A, 2, i, j), so = qb_load(A, 3, i, j);
      // slope_bar of each axis, and the sum the cell's own bar starts from
      sbs[0][t] = 0.5 * (e.r - w.r); sbs[1][t] = 0.5 * (e.u - w.u);
      sbs[2][t] = 0.5 * (e.v - w.v); sbs[3][t] = 0.5 * (e.p - w.p);
      sbs[4][t] = 0.5 * (n.r - so.r); sbs[5][t] = 0.5 * (n.u - so.u);
      sbs[6][t] = 0.5 * (n.v - so.v); sbs[7][t] = 0.5 * (n.p - so.p);
      Prim sum{0.0, 0.0, 0.0, 0.0};
      add(sum, e);
      add(sum, w);
      add(sum, n);
      add(sum, so);
      sbs[8][t] = sum.r; sbs[9][t] = sum.u; sbs[10][t] = sum.v; sbs[11][t] = sum.p;
    }
    __syncthreads();
  }
  const int li = threadIdx.x % CTX, lj = threadIdx.x / CTX;
  const int i = gi0 + li, j = gj0 + lj;
  if (i > g.nx + 1 || j > g.ny + 1) return;
  auto qat = [&](int ci, int cj) {
    const int t = (cj - gj0 + 2) * QX + (ci - gi0 + 2);
    return Prim{qs[0][t], qs[1][t], qs[2][t], qs[3][t]};
  };
  Prim qbar{0.0, 0.0, 0.0, 0.0};
  if (o2) {
    const int ot = (lj + 1) * BX + (li + 1);
    qbar = Prim{sbs[8][ot], sbs[9][ot], sbs[10][ot], sbs[11][ot]};
  } else {
    for (int side = 0; side < 4; ++side) add(qbar, qb_load(A, side, i, j));
  }
  if (o2) {
    for (int axis = 0; axis < 2; ++axis) {
      const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
      for (int o = -1; o <= 1; ++o) {
        const int ci = i + o * di, cj = j + o * dj;
        if (ci < -1 || cj < -1 || ci > g.nx || cj > g.ny) continue;  // no face states there
        const int bt = (cj - gj0 + 1) * BX + (ci - gi0 + 1);
        const Prim sb{sbs[4 * axis][bt], sbs[4 * axis + 1][bt], sbs[4 * axis + 2][bt],
                      sbs[4 * axis + 3][bt]};
        if (sb.r == 0.0 && sb.u == 0.0 && sb.v == 0.0 && sb.p == 0.0) continue;
        const Prim qm = qat(ci - di, cj - dj), qc = qat(ci, cj), qp = qat(ci + di, cj + dj);
        const double bk[4] = {qc.r - qm.r, qc.u - qm.u, qc.v - qm.v, qc.p - qm.p};
        const double fw[4] = {qp.r - qc.r, qp.u - qc.u, qp.v - qc.v, qp.p - qc.p};
        const double sbv[4] = {sb.r, sb.u, sb.v, sb.p};
        double contrib[4];
        for (int c = 0; c < 4; ++c) {
          double db, df;
          limited_slope_grad(bk[c], fw[c], g.limiter, db, df);
          double d = 0.0;
          if (o == 0) d = db - df;
          else if (o == 1) d = -db;
          else d = df;
          contrib[c] = sbv[c] * d;
        }
        qbar.r += contrib[0];
        qbar.u += contrib[1];
        qbar.v += contrib[2];
        qbar.p += contrib[3];
      }
    }
  }
  Cons ubar{0.0, 0.0, 0.0, 0.0};
  if (qbar.r != 0.0 || qbar.u != 0.0 || qbar.v != 0.0 || qbar.p != 0.0) {
    to_prim_vjp(adj_load_cell(A, i, j), g.gamma, qbar, ubar);
  }
  if (fold_interior(g, i, j)) {
    const long long k = (long long)j * A.rs + i;
    const double acc[4] = {0.0 + ubar.r, 0.0 + ubar.mx, 0.0 + ubar.my, 0.0 + ubar.E};
    for (int c = 0; c < 4; ++c) A.lam0[c * A.cs + k] = A.lam1[c * A.cs + k] + acc[c];
    return;
  }
  const long long t = gidx(g, i, j);
  A.ub[t] = ubar.r;
  A.ub[G + t] = ubar.mx;
  A.ub[2 * G + t] = ubar.my;
  A.ub[3 * G + t] = ubar.E;
}

// Where the explicit ghost rule reads grown cell (i,j) from: valid index and
// momentum flips, or false for a fixed state.
EC_HD bool ghost_source(const GridParams& g, int i, int j, int& ii, int& jj,
                                             bool& rx, bool& ry)
{
  if (i < 0 && g.bc[0] == BC_FIXED) return false;
  if (i >= g.nx && g.bc[1] == BC_FIXED) return false;
  if (j < 0 && g.bc[2] == BC_FIXED) return false;
  if (j >= g.ny && g.bc[3] == BC_FIXED) return false;
  ii = map_axis(i, g.nx, g.bc[0], g.bc[1], rx);
  jj = map_axis(j, g.ny, g.bc[2], g.bc[3], ry);
  return true;
}

// Fold of the ghost-cell bars onto valid cell (i,j) (the transpose of the
// ghost rule): lambda_n = lambda_next + the bars of every grown cell the
// explicit ghost rule reads from (i,j), with the reflected momentum negated.
EC_HD void adj_fold_cell(const AdjArgs& A, int i, int j)
{
  const GridParams& g = A.g;
  const long long G = grown_count(g);
  // grown columns / rows that can read (i,j): itself, then the ghosts in the
  // order -2, -1, n, n+1 (a ghost equal to i or j is the same candidate)
  const int ax[5] = {i, -2, -1, g.nx, g.nx + 1};
  const int ay[5] = {j, -2, -1, g.ny, g.ny + 1};
  int cx[5], cy[5], nxc = 0, nyc = 0;
  for (int a = 0; a < 5; ++a) {
    if (a > 0 && ax[a] == i) continue;
    bool r;
    if (map_axis(ax[a], g.nx, g.bc[0], g.bc[1], r) == i) cx[nxc++] = ax[a];
  }
  for (int b = 0; b < 5; ++b) {
    if (b > 0 && ay[b] == j) continue;
    bool r;
    if (map_axis(ay[b], g.ny, g.bc[2], g.bc[3], r) == j) cy[nyc++] = ay[b];
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < nyc; ++b) {
    for (int a = 0; a < nxc; ++a) {
      int ii, jj;
      bool rx, ry;
      if (!ghost_source(g, cx[a], cy[b], ii, jj, rx, ry) || ii != i || jj != j) continue;
      const long long k = gidx(g, cx[a], cy[b]);
      acc[0] += A.ub[k];
      acc[1] += rx ? -A.ub[G + k] : A.ub[G + k];
      acc[2] += ry ? -A.ub[2 * G + k] : A.ub[2 * G + k];
      acc[3] += A.ub[3 * G + k];
    }
  }
  const long long k = (long long)j * A.rs + i;
  for (int c = 0; c < 4; ++c) A.lam0[c * A.cs + k] = A.lam1[c * A.cs + k] + acc[c];
}

EC_HD void adj_fold_item(const AdjArgs& A, long long t)
{
  const int i = (int)(t % A.g.nx), j = (int)(t / A.g.nx);
  if (fold_interior(A.g, i, j)) return;  // written by the cell pass
  adj_fold_cell(A, i, j);
}

// The valid cells that are not fold-interior: rows 0, 1, ny-2, ny-1 and
// columns 0, 1, nx-2, nx-1 (all cells when the grid is narrower than 5).
EC_HD long long ring_count(const GridParams& g)
{
  if (g.nx < 5 || g.ny < 5) return (long long)g.nx * g.ny;
  return 4LL * g.nx + 4LL * (g.ny - 4);
}

EC_HD void ring_cell(const GridParams& g, long long t, int& i, int& j)
{
  if (g.nx < 5 || g.ny < 5) {
    i = (int)(t % g.nx);
    j = (int)(t / g.nx);
    return;
  }
  if (t < 4LL * g.nx) {
    const int r = (int)(t / g.nx);
    i = (int)(t % g.nx);
    j = r < 2 ? r : g.ny - 4 + r;
    return;
  }
  t -= 4LL * g.nx;
  const int c = (int)(t / (g.ny - 4));
  j = 2 + (int)(t % (g.ny - 4));
  i = c < 2 ? c : g.nx - 4 + c;
}

__global__ void __launch_bounds__(256) adj_fold_kernel(AdjArgs A)
{
  adj_priority();
  const long long n = ring_count(A.g);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j;
    ring_cell(A.g, t, i, j);
    if (!fold_interior(A.g, i, j)) adj_fold_cell(A, i, j);
  }
}

static int adj_blocks(long long n)
{
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  return b < 1 ? 1 : (int)b;
}

// The fold of one adjoint step and the prep pass of the next (earlier) one
// in a single launch: they touch disjoint data (the fold reads the bars and
// lambda of the later step, the prep writes the primitives of the earlier
// state), so the fold's latency hides behind the prep's work.
__global__ void __launch_bounds__(256) adj_fold_prep_kernel(AdjArgs F, AdjArgs P, int fold_blocks)
{
  adj_priority();
  if ((int)blockIdx.x < fold_blocks) {
    const long long n = ring_count(F.g);
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)fold_blocks * blockDim.x) {
      int i, j;
      ring_cell(F.g, t, i, j);
      if (!fold_interior(F.g, i, j)) adj_fold_cell(F, i, j);
    }
    return;
  }
  const long long G = grown_count(P.g);
  const long long nb = (long long)gridDim.x - fold_blocks;
  for (long long t = ((long long)blockIdx.x - fold_blocks) * blockDim.x + threadIdx.x; t < G;
       t += nb * blockDim.x) {
    adj_prep_item(P, t);
  }
}

static ec_status adjoint_args(ec_ctx* ctx, const double* U, const double* lam1, double* lam0,
                              long long rs, long long cs, double dt, AdjArgs* out)
{
  AdjWork* w = nullptr;
  ec_status st = adjoint_workspace(ctx, &w);
  if (st != EC_OK) return st;
  *out = AdjArgs{U, lam1, lam0, rs, cs, ctx->grid, dt, w->qb, w->ub,
                 ctx->grid.riemann == 3 ? w->sel : nullptr, w->mbar, w->q};
  return EC_OK;
}

// Face, ratio and cell passes of one step (the prep pass must have run).
static ec_status adjoint_passes(ec_ctx* ctx, const AdjArgs& A)
{
  const GridParams& g = A.g;
  const dim3 fgrid(adj_blocks(face_items(g)));
  switch (g.riemann) {
    case 0: hipLaunchKernelGGL(adj_faces_kernel<0>, fgrid, dim3(256), 0, ctx->stream, A); break;
    case 1: hipLaunchKernelGGL(adj_faces_kernel<1>, fgrid, dim3(256), 0, ctx->stream, A); break;
    case 2: hipLaunchKernelGGL(adj_faces_kernel<2>, fgrid, dim3(256), 0, ctx->stream, A); break;
    default: hipLaunchKernelGGL(adj_faces_kernel<3>, fgrid, dim3(256), 0, ctx->stream, A); break;
  }
  if (g.riemann == 3) {
    const long long nr = (long long)(g.nx + 1) * (g.ny + 2) + (long long)(g.nx + 2) * (g.ny + 1);
    hipLaunchKernelGGL(adj_ratios_kernel, dim3(adj_blocks(nr)), dim3(256), 0, ctx->stream, A);
  }
  {
    const int tx = (g.nx + 4 + CTX - 1) / CTX, ty = (g.ny + 4 + CTY - 1) / CTY;
    hipLaunchKernelGGL(adj_cells_tile_kernel, dim3(tx * ty), dim3(CNT), 0, ctx->stream, A);
  }
  EC_HIP_TRY(ctx, hipGetLastError());
  return EC_OK;
}

ec_status launch_adjoint_step(ec_ctx* ctx, const double* U, const double* lam1, double* lam0,
                              long long rs, long long cs, double dt)
{
  AdjArgs A;
  ec_status st = adjoint_args(ctx, U, lam1, lam0, rs, cs, dt, &A);
  if (st != EC_OK) return st;
  hipLaunchKernelGGL(adj_prep_kernel, dim3(adj_blocks(grown_count(A.g))), dim3(256), 0, ctx->stream, A);
  if ((st = adjoint_passes(ctx, A)) != EC_OK) return st;
  hipLaunchKernelGGL(adj_fold_kernel, dim3(adj_blocks(ring_count(A.g))), dim3(256), 0, ctx->stream, A);
  EC_HIP_TRY(ctx, hipGetLastError());
  return EC_OK;
}

}  // namespace ec

extern "C" ec_status ec_adjoint_step_host(const ec_grid_desc* desc, const double* U_n,
                                          const double* lambda_next, double* lambda_n, double dt)
{
  if (!desc || !U_n || !lambda_next || !lambda_n) return EC_ERR_INVALID_ARGUMENT;
  const GridParams g = make_grid_params(*desc);
  const long long G = (long long)(g.nx + 4) * (g.ny + 4);
  const long long nf = (long long)(g.nx + 1) * g.ny + (long long)g.nx * (g.ny + 1);
  const long long nr = (long long)(g.nx + 1) * (g.ny + 2) + (long long)(g.nx + 2) * (g.ny + 1);
  // NaN-filled scratch: a pass that read a slot no earlier pass wrote would
  // show up in the result.
  const double nan = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> qb(16 * G, nan), ub(4 * G, nan), mbar(nf, nan), q(12 * G, nan);
  std::vector<int> sel(nf, -1);
  const long long n = (long long)g.nx * g.ny;
  AdjArgs A{U_n, lambda_next, lambda_n, g.nx, n, g, dt, qb.data(), ub.data(),
            g.riemann == 3 ? sel.data() : nullptr, mbar.data(), q.data()};
  for (long long t = 0; t < G; ++t) adj_prep_item(A, t);
  for (long long t = 0; t < face_items(g); ++t) adj_face_item(A, t);
  if (g.riemann == 3) {
    for (long long t = 0; t < nr; ++t) adj_ratio_item(A, t);
  }
  for (long long t = 0; t < G; ++t) adj_cell_item(A, t);
  for (long long t = 0; t < n; ++t) adj_fold_item(A, t);
  return EC_OK;
}

extern "C" ec_status ec_adjoint_step(ec_ctx* ctx, const double* U_n, const double* lambda_next,
                                     double* lambda_n, const ec_layout* lay, double dt)
{
  if (!ctx || !U_n || !lambda_next || !lambda_n || !lay || lambda_n == lambda_next ||
      lay->row_stride < ctx->grid.nx || lay->comp_stride < lay->row_stride * ctx->grid.ny ||
      !(dt >= 0.0)) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (has_exact_side(ctx->grid)) {
    ctx->last_error = "the adjoint does not take exact-Dirichlet sides";
    return EC_ERR_UNSUPPORTED;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  return launch_adjoint_step(ctx, U_n, lambda_next, lambda_n, lay->row_stride, lay->comp_stride, dt);
}

// ============================================================ checkpointing
namespace ec {
ec_status launch_explicit_step(ec_ctx* ctx, const double* uin, double* uout, long long rs,
                               long long cs, double dt, int device_dt);
ec_status launch_explicit_step_on(ec_ctx* ctx, hipStream_t stream, const double* uin, double* uout,
                                  long long rs, long long cs, double dt, int device_dt);
ec_status launch_state_reduce(ec_ctx* ctx, const double* u, long long rs, long long cs, int mode);

// Copy a state between layouts (4 planes of nx x ny).
static ec_status copy_state(ec_ctx* ctx, double* dst, long long drs, long long dcs, const double* src,
                            long long srs, long long scs)
{
  const size_t nx = (size_t)ctx->grid.nx, ny = (size_t)ctx->grid.ny;
  for (int c = 0; c < 4; ++c) {
    EC_HIP_TRY(ctx, hipMemcpy2DAsync(dst + c * dcs, drs * sizeof(double), src + c * scs,
                                     srs * sizeof(double), nx * sizeof(double), ny,
                                     hipMemcpyDeviceToDevice, ctx->stream));
  }
  return EC_OK;
}

static ec_status ensure_states(ec_ctx* ctx, AdjWork* w, int count)
{
  if (w->seg_cap >= count) return EC_OK;
  if (w->states) (void)hipFree(w->states);
  w->states = nullptr;
  w->seg_cap = 0;
  const size_t bytes = (size_t)count * 4 * (size_t)w->nx * w->ny * sizeof(double);
  if (hipMalloc((void**)&w->states, bytes) != hipSuccess) return EC_ERR_OUT_OF_MEMORY;
  w->seg_cap = count;
  (void)ctx;
  return EC_OK;
}

static ec_status ensure_lam(AdjWork* w)
{
  if (w->lam) return EC_OK;
  const size_t bytes = 2 * 4 * (size_t)w->nx * w->ny * sizeof(double);
  return hipMalloc((void**)&w->lam, bytes) == hipSuccess ? EC_OK : EC_ERR_OUT_OF_MEMORY;
}

}  // namespace ec

static bool lay_ok(const ec_ctx* ctx, const ec_layout* l)
{
  return l && l->row_stride >= ctx->grid.nx && l->comp_stride >= l->row_stride * ctx->grid.ny;
}

extern "C" ec_status ec_forward_checkpointed(ec_ctx* ctx, const double* U0, double* U_final,
                                             const ec_layout* lay, int nsteps, int checkpoint_every,
                                             double cfl, double stop_time, double* checkpoints,
                                             double* dts, int* steps_done)
{
  if (!ctx || !U0 || !U_final || !lay_ok(ctx, lay) || nsteps < 1 || checkpoint_every < 1 ||
      !checkpoints || !dts || !steps_done) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (has_exact_side(ctx->grid)) {
    ctx->last_error = "the adjoint does not take exact-Dirichlet sides";
    return EC_ERR_UNSUPPORTED;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  AdjWork* w = nullptr;
  ec_status st = adjoint_workspace(ctx, &w);
  if (st != EC_OK) return st;
  if ((st = ensure_states(ctx, w, 2)) != EC_OK) return st;
  if (w->dt_cap < nsteps) {
    if (w->dtlog) (void)hipFree(w->dtlog);
    w->dtlog = nullptr;
    w->dt_cap = 0;
    if (hipMalloc((void**)&w->dtlog, (size_t)nsteps * sizeof(double)) != hipSuccess) {
      return EC_ERR_OUT_OF_MEMORY;
    }
    w->dt_cap = nsteps;
  }
  const long long nx = ctx->grid.nx, n = nx * ctx->grid.ny;
  double* A = w->states;
  double* B = w->states + 4 * n;
  if ((st = copy_state(ctx, A, nx, n, U0, lay->row_stride, lay->comp_stride)) != EC_OK) return st;
  DevCtrl& h = *ctx->ctrl_host;
  std::memset(&h, 0, sizeof(DevCtrl));
  h.stop_time = stop_time;
  h.cfl = cfl;
  h.dt_min = INFINITY;
  h.max_step = nsteps;
  EC_HIP_TRY(ctx, hipMemsetAsync(w->dtlog, 0, (size_t)nsteps * sizeof(double), ctx->stream));
  EC_HIP_TRY(ctx, hipMemcpyAsync(ctx->ctrl, &h, sizeof(DevCtrl), hipMemcpyHostToDevice, ctx->stream));
  if ((st = launch_state_reduce(ctx, A, nx, n, 3)) != EC_OK) return st;
  ctx->dt_log = w->dtlog;
  for (int s = 0; s < nsteps; ++s) {
    if (s % checkpoint_every == 0) {
      EC_HIP_TRY(ctx, hipMemcpyAsync(checkpoints + (size_t)(s / checkpoint_every) * 4 * n, A,
                                     4 * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    st = launch_explicit_step(ctx, A, B, nx, n, 0.0, 1);
    if (st != EC_OK) {
      ctx->dt_log = nullptr;
      return st;
    }
    double* t = A;
    A = B;
    B = t;
  }
  ctx->dt_log = nullptr;
  if ((st = copy_state(ctx, U_final, lay->row_stride, lay->comp_stride, A, nx, n)) != EC_OK) return st;
  EC_HIP_TRY(ctx, hipMemcpyAsync(dts, w->dtlog, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream));
  EC_HIP_TRY(ctx, hipMemcpyAsync(&h, ctx->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, ctx->stream));
  EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *steps_done = h.step;
  return EC_OK;
}

extern "C" ec_status ec_adjoint_checkpointed(ec_ctx* ctx, const double* checkpoints, const ec_layout* lay,
                                             int nsteps, int checkpoint_every, const double* dts,
                                             const double* lambda_N, double* lambda_0)
{
  if (!ctx || !checkpoints || !lay_ok(ctx, lay) || nsteps < 1 || checkpoint_every < 1 || !dts ||
      !lambda_N || !lambda_0) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (has_exact_side(ctx->grid)) {
    ctx->last_error = "the adjoint does not take exact-Dirichlet sides";
    return EC_ERR_UNSUPPORTED;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  AdjWork* w = nullptr;
  ec_status st = adjoint_workspace(ctx, &w);
  if (st != EC_OK) return st;
  const int K = checkpoint_every;
  // Two segment buffers: segment seg - 1 is recomputed from its checkpoint on
  // a second stream while the adjoint steps of segment seg run on the
  // context's stream (the grids of one step fill only part of the chip).
  if ((st = ensure_states(ctx, w, 2 * (K < 2 ? 2 : K))) != EC_OK) return st;
  if ((st = ensure_lam(w)) != EC_OK) return st;
  if (!w->rstream) {
    EC_HIP_TRY(ctx, hipStreamCreateWithFlags(&w->rstream, hipStreamNonBlocking));
    EC_HIP_TRY(ctx, hipEventCreateWithFlags(&w->ev_start, hipEventDisableTiming));
    for (int b = 0; b < 2; ++b) {
      EC_HIP_TRY(ctx, hipEventCreateWithFlags(&w->ev_rec[b], hipEventDisableTiming));
      EC_HIP_TRY(ctx, hipEventCreateWithFlags(&w->ev_use[b], hipEventDisableTiming));
    }
  }
  const long long nx = ctx->grid.nx, n = nx * ctx->grid.ny;
  const int slots = K < 2 ? 2 : K;
  double* la = w->lam;
  double* lb = w->lam + 4 * n;
  if ((st = copy_state(ctx, la, nx, n, lambda_N, lay->row_stride, lay->comp_stride)) != EC_OK) return st;
  // the recompute stream starts after the work already queued on the context's stream
  EC_HIP_TRY(ctx, hipEventRecord(w->ev_start, ctx->stream));
  EC_HIP_TRY(ctx, hipStreamWaitEvent(w->rstream, w->ev_start, 0));
  const int nseg = (nsteps + K - 1) / K;
  bool used[2] = {false, false};  // ev_use[b] recorded (buffer b read by an adjoint segment)
  auto seg_states = [&](int seg) { return w->states + (size_t)(seg & 1) * slots * 4 * n; };
  auto recompute = [&](int seg) -> ec_status {
    const int b = seg & 1, start = seg * K;
    const int len = nsteps - start < K ? nsteps - start : K;
    double* S = seg_states(seg);
    if (used[b]) EC_HIP_TRY(ctx, hipStreamWaitEvent(w->rstream, w->ev_use[b], 0));
    EC_HIP_TRY(ctx, hipMemcpyAsync(S, checkpoints + (size_t)seg * 4 * n, 4 * n * sizeof(double),
                                   hipMemcpyDeviceToDevice, w->rstream));
    for (int s = 1; s < len; ++s) {
      const ec_status e = launch_explicit_step_on(ctx, w->rstream, S + (size_t)(s - 1) * 4 * n,
                                                  S + (size_t)s * 4 * n, nx, n, dts[start + s - 1], 2);
      if (e != EC_OK) return e;
    }
    EC_HIP_TRY(ctx, hipEventRecord(w->ev_rec[b], w->rstream));
    return EC_OK;
  };
  AdjArgs prev{};
  bool pending = false;  // a step whose fold has not been launched yet
  if ((st = recompute(nseg - 1)) != EC_OK) return st;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int b = seg & 1, start = seg * K;
    const int len = nsteps - start < K ? nsteps - start : K;
    if (seg > 0 && (st = recompute(seg - 1)) != EC_OK) return st;
    EC_HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, w->ev_rec[b], 0));
    double* S = seg_states(seg);
    for (int s = len - 1; s >= 0; --s) {
      AdjArgs A;
      if ((st = adjoint_args(ctx, S + (size_t)s * 4 * n, la, lb, nx, n, dts[start + s], &A)) != EC_OK) {
        return st;
      }
      const long long G = grown_count(A.g);
      if (pending) {
        // the previous step's fold shares a launch with this step's prep
        const int fb = adj_blocks(ring_count(A.g));
        hipLaunchKernelGGL(adj_fold_prep_kernel, dim3(fb + adj_blocks(G)), dim3(256), 0, ctx->stream,
                           prev, A, fb);
      } else {
        hipLaunchKernelGGL(adj_prep_kernel, dim3(adj_blocks(G)), dim3(256), 0, ctx->stream, A);
      }
      if ((st = adjoint_passes(ctx, A)) != EC_OK) return st;
      prev = A;
      pending = true;
      double* t = la;
      la = lb;
      lb = t;
    }
    // the buffer is free once this segment's last reads (its step 0 prep,
    // already enqueued) are done; the pending fold reads lambda only
    EC_HIP_TRY(ctx, hipEventRecord(w->ev_use[b], ctx->stream));
    used[b] = true;
  }
  if (pending) {
    hipLaunchKernelGGL(adj_fold_kernel, dim3(adj_blocks(ring_count(prev.g))), dim3(256), 0, ctx->stream,
                       prev);
    EC_HIP_TRY(ctx, hipGetLastError());
  }
  if ((st = copy_state(ctx, lambda_0, lay->row_stride, lay->comp_stride, la, nx, n)) != EC_OK) return st;
  EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return EC_OK;
}
