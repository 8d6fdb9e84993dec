// ms_physics.hip -- assembly, Rhie-Chow, correction and diagnostics kernels
// plus the SIMPLE / PISO drivers, all resident on one gfx950 device.
//
// Every cell-based kernel is an atomics-free gather: one thread per cell walks
// its faces in ascending face order (the per-cell accumulation order of the
// reference's face loops, interior faces before boundary faces where the
// reference uses two loops), so assembled coefficients, right-hand sides and
// gradients are bit-identical to the reference (no FMA contraction; every
// expression keeps the reference's evaluation order).  Face kernels compute
// face values without scatter.
//
// Reference routines (src/unstructured/...):
//   k_gradients      mod_uns_fields.f90:164-258, grad_scalar :355-374
//   k_limiter        mod_uns_simple.f90:1346-1408 (+ temperature :2302-2361)
//   k_momentum       mod_uns_simple.f90:928-1338
//   k_energy         mod_uns_simple.f90:1748-2037
//   k_rhiechow       mod_uns_simple.f90:1413-1450
//   outflow kernels  mod_uns_simple.f90:1463-1531
//   k_ppe_af/k_ppe   mod_uns_simple.f90:1552-1644, mod_uns_simple_mpi.f90:882-968
//   k_correct_*      mod_uns_simple.f90:1649-1739
//   k_diagnostics    mod_uns_simple_mpi.f90:502-526, nan_guard_mpi :1024-1074
//   k_wall_forces    compute_forces, mod_uns_forces.f90:135-156
//   k_wall_samples   owner values of write_cp_distribution (:238-246) and
//                    compute_separation (:320-333); k_wake_end (:365-370)
//   drivers          simple_run (mod_uns_simple.f90:88-288),
//                    piso_run_mpi (mod_uns_simple_mpi.f90:298-571)
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "ms_reduce.hpp"
#include "ms_system.hpp"

namespace ms {
namespace {

struct Ramp {
  double inlet, pval;
};

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double ddot3(const double *a, const double *b) {
  double s = 0.0;
  s = s + a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}
// flang runtime NORM2 order (see norm2_ref)
__device__ __forceinline__ double dnorm3(const double *v) {
  double mx = 0.0, s = 0.0;
  for (int i = 0; i < 3; i++) {
    double ax = fabs(v[i]);
    if (mx == 0.0) {
      mx = ax;
    } else if (ax > mx) {
      double t = mx / ax;
      double tsq = t * t;
      s = s * tsq;
      s = s + tsq;
      mx = ax;
    } else {
      double t = ax / mx;
      s = s + t * t;
    }
  }
  return mx * sqrt(1.0 + s);
}

// bc_face_vel, mod_uns_bc.f90:400-509.  sf/area equals sf/norm2(sf) exactly
// because area is norm2 of the same vector up to sign.
__device__ void bc_face_vel(const DMesh &m, int f, const double *uP, double *uf, const Ramp &rp) {
  const DevBC &g = m.bc[m.fgrp[f] - 1];
  const double *xf = &m.xf[3 * (size_t)f], *sf = &m.sf[3 * (size_t)f];
  const double a = m.area[f];
  double nv[3], un;
  switch (g.btype) {
    case MS_BC_WALL:
      if (g.has_lid && fabs(xf[g.lid_dir - 1] - g.lid_coord) <= g.lid_tol) {
        uf[0] = g.lid_vel[0]; uf[1] = g.lid_vel[1]; uf[2] = g.lid_vel[2];
      } else {
        uf[0] = uf[1] = uf[2] = 0.0;
      }
      break;
    case MS_BC_SYMMETRY:
      for (int d = 0; d < 3; d++) nv[d] = sf[d] / a;
      un = ddot3(uP, nv);
      for (int d = 0; d < 3; d++) uf[d] = uP[d] - 2.0 * un * nv[d];
      break;
    case MS_BC_VINLET:
      uf[0] = g.uvel[0]; uf[1] = g.uvel[1]; uf[2] = g.uvel[2];
      break;
    case MS_BC_VINLET_PARAB: {
      for (int d = 0; d < 3; d++) nv[d] = sf[d] / a;
      double xi = (xf[g.prfl_axis - 1] - g.prfl_origin) / g.prfl_span;
      xi = dmax(0.0, dmin(1.0, xi));
      double uspc = 6.0 * g.prfl_mean * xi * (1.0 - xi);
      for (int d = 0; d < 3; d++) uf[d] = -uspc * nv[d];
    } break;
    case MS_BC_MASSINLET:
      for (int d = 0; d < 3; d++) nv[d] = sf[d] / a;
      for (int d = 0; d < 3; d++) uf[d] = -g.uspeed * rp.inlet * nv[d];
      break;
    case MS_BC_FARFIELD:
      for (int d = 0; d < 3; d++) nv[d] = sf[d] / a;
      un = ddot3(g.uvel, nv);
      if (un < 0.0) {
        uf[0] = g.uvel[0]; uf[1] = g.uvel[1]; uf[2] = g.uvel[2];
      } else {
        uf[0] = uP[0]; uf[1] = uP[1]; uf[2] = uP[2];
      }
      break;
    case MS_BC_SLIPWALL:
      for (int d = 0; d < 3; d++) nv[d] = sf[d] / a;
      un = ddot3(uP, nv);
      for (int d = 0; d < 3; d++) uf[d] = uP[d] - un * nv[d];
      break;
    case MS_BC_INTERFACE:   // :460-474: extrapolated under the interface pressure model
      if (m.iface_pdir) {
        uf[0] = uP[0]; uf[1] = uP[1]; uf[2] = uP[2];
      } else {
        const double *v = &m.iface_vel[3 * (size_t)f];
        uf[0] = v[0]; uf[1] = v[1]; uf[2] = v[2];
      }
      break;
    default:
      uf[0] = uP[0]; uf[1] = uP[1]; uf[2] = uP[2];
  }
}

__device__ __forceinline__ double bc_face_p(const DMesh &m, int f, double pP, const Ramp &rp) {
  const DevBC &g = m.bc[m.fgrp[f] - 1];
  if (g.btype == MS_BC_POUTLET || g.btype == MS_BC_FARFIELD) return g.pval * rp.pval;
  if (g.btype == MS_BC_INTERFACE && m.iface_pdir) return m.iface_p[f];
  return pP;
}

// bc_face_p_dirichlet (mod_uns_bc.f90:879-893): pressure outlets, far field,
// and interface faces under the interface pressure model
__device__ __forceinline__ bool p_dirichlet(const DMesh &m, int f) {
  int gi = m.fgrp[f];
  if (gi <= 0) return false;
  int bt = m.bc[gi - 1].btype;
  return bt == MS_BC_POUTLET || bt == MS_BC_FARFIELD || (bt == MS_BC_INTERFACE && m.iface_pdir);
}

// the PPE's pressure-Dirichlet test: the serial ppe_assembly uses
// bc_face_p_dirichlet (mod_uns_simple.f90:1585, :1624), ppe_assembly_mpi the
// types alone (mod_uns_simple_mpi.f90:903-907, :942-946)
__device__ __forceinline__ bool p_dirichlet_ppe(const DMesh &m, int f) {
  int gi = m.fgrp[f];
  if (gi <= 0) return false;
  int bt = m.bc[gi - 1].btype;
  return bt == MS_BC_POUTLET || bt == MS_BC_FARFIELD || (bt == MS_BC_INTERFACE && m.iface_ppe);
}

// bc_face_T (static part precomputed per face)
__device__ __forceinline__ double bc_face_T(const DMesh &m, int f, double Tc) {
  return m.tmode[f] == 1 ? m.tval[f] : Tc;
}

// rr = xf[f] - xc[cell] for the cell on side `far` (0: c0, 1: c1) of face f;
// the c1 cell of a joined periodic face sees the face at xf - shift.
__device__ __forceinline__ void face_offset(const DMesh &m, int f, int cell, bool far, double *rr) {
  const double *xf = &m.xf[3 * (size_t)f], *xc = &m.xc[3 * (size_t)cell];
  const int k = (far && m.nper) ? m.pidx[f] : 0;
  if (k == 0) {
    for (int d = 0; d < 3; d++) rr[d] = xf[d] - xc[d];
  } else {
    for (int d = 0; d < 3; d++) rr[d] = (xf[d] - m.psh[k - 1][d]) - xc[d];
  }
}

// kink_face_pressure (mod_uns_fields.f90:328-349): one-sided quadratic
// reconstruction at a fluid/porous face from the cell gradients gp
__device__ __forceinline__ double kink_pf(const DMesh &m, const double *p, const double *gp, int f, int a, int b) {
  double dvec[3], nhat[3];
  const int k = m.nper ? m.pidx[f] : 0;
  for (int d = 0; d < 3; d++)
    dvec[d] = (k ? m.xc[3 * (size_t)b + d] + m.psh[k - 1][d] : m.xc[3 * (size_t)b + d]) - m.xc[3 * (size_t)a + d];
  const double dn = m.dn[f];
  for (int d = 0; d < 3; d++) nhat[d] = dvec[d] / dn;
  const double l = m.lf[f];
  const double d0 = (1.0 - l) * dn, d1 = -l * dn;
  return 0.5 * (p[a] + d0 * ddot3(&gp[3 * (size_t)a], nhat) + p[b] + d1 * ddot3(&gp[3 * (size_t)b], nhat));
}

__device__ __forceinline__ int face_flags(const DMesh &m, int f) { return m.has_face_flags ? m.fflag[f] : 0; }
__device__ __forceinline__ bool solid_cell(const DFields &F, int r) { return F.cztype && F.cztype[r] == CZ_SOLID; }

// Row of this thread in the cell-wise assembly kernels (-1: none).  On a
// two-colour numbering each workgroup takes kBlock/2 rows of colour 0 and
// the kBlock/2 rows of colour 1 at the same position of their colour's block
// (the same aggregate, the same place in it), so most of a cell's neighbours
// are rows of its own workgroup and its neighbour gathers hit the cache; the
// workgroups are spread over the XCDs in contiguous ranges.  Each row's
// arithmetic does not depend on the mapping.
__device__ __forceinline__ int cell_row(const DMesh &m) {
  const int b = xcd_block(blockIdx.x, gridDim.x), t = threadIdx.x;
  if (m.pair_n0 > 0) {
    constexpr int h = kBlock / 2;
    if (t < h) {
      const int i = b * h + t;
      return i < m.pair_n0 ? i : -1;
    }
    const int i = b * h + (t - h);
    return i < m.nc - m.pair_n0 ? m.pair_n0 + i : -1;
  }
  const int r = b * kBlock + t;
  return r < m.nc ? r : -1;
}

// ------------------------------------------------------------- gradients
__global__ void k_gradients(DMesh m, DFields F, Ramp rp, int do_T) {
  const int r = cell_row(m);
  if (r < 0) return;
  double gp[3] = {0, 0, 0}, gu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gt[3] = {0, 0, 0}, gs[3] = {0, 0, 0};
  const double *uP = &F.u[3 * (size_t)r];
  for (int k = m.cf_ptr[r]; k < m.cf_ptr[r + 1]; k++) {
    const int e = m.cf_ent[k], f = e >> 1, side = e & 1;
    const int o = m.e_oth[k];
    const int a = side == 0 ? r : o, b = side == 0 ? o : r;   // c0 and c1 of the face (b < 0 on the boundary)
    const double *sf = &m.e_sf[3 * (size_t)k];
    double phip, phiu[3], phit = 0.0, phis = 0.0;
    if (b >= 0) {
      const double l = m.e_lf[k];
      if (face_flags(m, f) & FF_KINK) phip = kink_pf(m, F.p, F.gp_prev, f, a, b);
      else phip = l * F.p[a] + (1.0 - l) * F.p[b];
      for (int c = 0; c < 3; c++) phiu[c] = l * F.u[3 * (size_t)a + c] + (1.0 - l) * F.u[3 * (size_t)b + c];
      if (do_T) {
        phit = l * F.T[a] + (1.0 - l) * F.T[b];
        phis = l * F.T_s[a] + (1.0 - l) * F.T_s[b];
      }
    } else {
      phip = bc_face_p(m, f, F.p[r], rp);
      bc_face_vel(m, f, uP, phiu, rp);
      if (do_T) {
        phit = bc_face_T(m, f, F.T[r]);
        phis = bc_face_T(m, f, F.T_s[r]);
      }
    }
    if (side == 0) {
      for (int d = 0; d < 3; d++) {
        gp[d] = gp[d] + phip * sf[d];
        for (int c = 0; c < 3; c++) gu[c + 3 * d] = gu[c + 3 * d] + phiu[c] * sf[d];
        gt[d] = gt[d] + phit * sf[d];
        gs[d] = gs[d] + phis * sf[d];
      }
    } else {
      for (int d = 0; d < 3; d++) {
        gp[d] = gp[d] - phip * sf[d];
        for (int c = 0; c < 3; c++) gu[c + 3 * d] = gu[c + 3 * d] - phiu[c] * sf[d];
        gt[d] = gt[d] - phit * sf[d];
        gs[d] = gs[d] - phis * sf[d];
      }
    }
  }
  const double v = m.vol[r];
  for (int d = 0; d < 3; d++) F.gp[3 * (size_t)r + d] = gp[d] / v;
  for (int j = 0; j < 9; j++) F.gu[9 * (size_t)r + j] = gu[j] / v;
  if (do_T)
    for (int d = 0; d < 3; d++) {
      F.gt[3 * (size_t)r + d] = gt[d] / v;
      F.gts[3 * (size_t)r + d] = gs[d] / v;
    }
}

// ------------------------------------------------ Barth-Jespersen limiter
// comp 0..2: velocity component (velocity_limiter); comp 3: temperature.
__global__ void k_limiter(DMesh m, DFields F, Ramp rp, int comp, double *psi) {
  const int r = cell_row(m);
  if (r < 0) return;
  const bool isT = comp == 3;
  const double uc = isT ? F.T[r] : F.u[3 * (size_t)r + comp];
  double pmin = uc, pmax = uc;
  const int k0 = m.cf_ptr[r], k1 = m.cf_ptr[r + 1];
  for (int k = k0; k < k1; k++) {
    const int f = m.cf_ent[k] >> 1, o = m.e_oth[k];
    double v;
    if (o >= 0) {
      v = isT ? F.T[o] : F.u[3 * (size_t)o + comp];
    } else if (isT) {
      v = bc_face_T(m, f, F.T[r]);
    } else {
      double uf[3];
      bc_face_vel(m, f, &F.u[3 * (size_t)r], uf, rp);
      v = uf[comp];
    }
    pmin = dmin(pmin, v);
    pmax = dmax(pmax, v);
  }
  double ps = 1.0, g[3];
  for (int d = 0; d < 3; d++) g[d] = isT ? F.gt[3 * (size_t)r + d] : F.gu[9 * (size_t)r + comp + 3 * d];
  for (int k = k0; k < k1; k++) {
    const int e = m.cf_ent[k];
    double rr[3];
    face_offset(m, e >> 1, r, (e & 1) != 0, rr);
    const double fr = uc + ddot3(g, rr);
    const double df = fr - uc;
    if (df > 0.0) ps = dmin(ps, (pmax - uc) / df);
    else if (df < 0.0) ps = dmin(ps, (pmin - uc) / df);
  }
  psi[r] = dmax(0.0, dmin(1.0, ps));
}

// The three velocity components' limiters in one pass: psi[c * n + r].
// Equal to three k_limiter launches when no boundary face value of one
// component depends on another component (no symmetry or slip-wall group),
// so that computing the limiters of components 1 and 2 before the solve of
// component 0 changes nothing.
__global__ void k_limiter3(DMesh m, DFields F, Ramp rp, double *psi) {
  const int r = cell_row(m);
  if (r < 0) return;
  const double *uP = &F.u[3 * (size_t)r];
  double uc[3], pmin[3], pmax[3];
  for (int c = 0; c < 3; c++) uc[c] = pmin[c] = pmax[c] = uP[c];
  const int k0 = m.cf_ptr[r], k1 = m.cf_ptr[r + 1];
  for (int k = k0; k < k1; k++) {
    const int o = m.e_oth[k];
    double v[3];
    if (o >= 0) {
      for (int c = 0; c < 3; c++) v[c] = F.u[3 * (size_t)o + c];
    } else {
      bc_face_vel(m, m.cf_ent[k] >> 1, uP, v, rp);
    }
    for (int c = 0; c < 3; c++) {
      pmin[c] = dmin(pmin[c], v[c]);
      pmax[c] = dmax(pmax[c], v[c]);
    }
  }
  double ps[3] = {1.0, 1.0, 1.0}, g[3][3];
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 3; d++) g[c][d] = F.gu[9 * (size_t)r + c + 3 * d];
  for (int k = k0; k < k1; k++) {
    const int e = m.cf_ent[k];
    double rr[3];
    face_offset(m, e >> 1, r, (e & 1) != 0, rr);
    for (int c = 0; c < 3; c++) {
      const double fr = uc[c] + ddot3(g[c], rr);
      const double df = fr - uc[c];
      if (df > 0.0) ps[c] = dmin(ps[c], (pmax[c] - uc[c]) / df);
      else if (df < 0.0) ps[c] = dmin(ps[c], (pmin[c] - uc[c]) / df);
    }
  }
  for (int c = 0; c < 3; c++) psi[(size_t)c * m.nc + r] = dmax(0.0, dmin(1.0, ps[c]));
}

// --------------------------------------------------------------- momentum
struct MomParams {
  int comp, ts_order, pimple, boussinesq, body_on;
  double rho, mu, alpha_u, dt, beta, tref, blend, nonorth, grav, body;
};

// The boundary faces, sources, time terms and the diagonal of one
// component (:1123-1338), continuing ap and rhs from the interior faces.
__device__ __forceinline__ void mom_tail(const DMesh &m, const DFields &F, const Ramp &rp, const MomParams &q, int r,
                                         int k0, int k1, double ap, double rhs, double *val, double *rhs_out,
                                         double *ap_out) {
  const int comp = q.comp;
  // boundary faces (:1123-1227)
  const double *uP = &F.u[3 * (size_t)r];
  for (int k = k0; k < k1; k++) {
    if (m.e_oth[k] >= 0) continue;
    const int f = m.cf_ent[k] >> 1;
    double uf[3];
    bc_face_vel(m, f, uP, uf, rp);
    const double pf = bc_face_p(m, f, F.p[r], rp);
    const double D = m.Dmb[f];
    const double Ff = q.rho * ddot3(uf, &m.sf[3 * (size_t)f]);
    switch (m.bc[m.fgrp[f] - 1].btype) {
      case MS_BC_WALL: case MS_BC_SYMMETRY: case MS_BC_VINLET: case MS_BC_VINLET_PARAB:
      case MS_BC_SLIPWALL: case MS_BC_MASSINLET:
        ap = ap + D + dmax(Ff, 0.0);
        rhs = rhs + D * uf[comp] - Ff * uf[comp];
        break;
      case MS_BC_POUTLET: case MS_BC_OUTFLOW:
        ap = ap + dmax(Ff, 0.0);
        rhs = rhs - dmin(Ff, 0.0) * uP[comp];
        break;
      case MS_BC_FARFIELD:
        if (Ff < 0.0) {
          ap = ap + D + dmax(Ff, 0.0);
          rhs = rhs + D * uf[comp] - Ff * uf[comp];
        } else {
          ap = ap + Ff;
        }
        break;
      case MS_BC_INTERFACE:   // :1144-1206
        if (!m.iface_pdir) {
          ap = ap + D + dmax(Ff, 0.0);
          rhs = m.iface_zhang ? rhs + D * uf[comp] - dmin(Ff, 0.0) * uf[comp] : rhs + D * uf[comp] - Ff * uf[comp];
        } else {
          ap = ap + dmax(Ff, 0.0);
          rhs = rhs - dmin(Ff, 0.0) * uP[comp];
          if (m.iface_slip) {
            const double ct = m.iface_slip_C[f];
            if (ct > 0.0) {
              double nh2[3];
              for (int d = 0; d < 3; d++) nh2[d] = m.sf[3 * (size_t)f + d] / m.area[f];
              const double *pv = &m.iface_peer_vel[3 * (size_t)f];
              const double un_own = ddot3(uP, nh2), un_peer = ddot3(pv, nh2);
              ap = ap + ct * (1.0 - nh2[comp] * nh2[comp]);
              rhs = rhs + ct * (pv[comp] - un_peer * nh2[comp]);
              rhs = rhs + ct * nh2[comp] * (un_own - uP[comp] * nh2[comp]);
            }
          }
        }
        break;
      default:
        break;
    }
    rhs = rhs - pf * m.sf[3 * (size_t)f + comp];
  }
  const double vol = m.vol[r];
  if (q.boussinesq) rhs = rhs - q.rho * q.beta * (F.T[r] - q.tref) * q.grav * vol;
  if (q.body_on) rhs = rhs + q.body * vol;
  const double K = F.perm_dir[3 * (size_t)r + comp];
  if (K > 0.0) ap = ap + q.mu / K * vol;
  if (F.inertial[r] > 0.0) {
    const double um = sqrt(uP[0] * uP[0] + uP[1] * uP[1] + uP[2] * uP[2]);
    rhs = rhs - q.rho * F.inertial[r] * um * uP[comp] * vol;
  }
  if (solid_cell(F, r)) ap = ap + 1.0e15 * vol;   // stationary solid (:1282-1283)
  if (q.ts_order == 1) {
    const double at = q.rho * vol / q.dt;
    ap = ap + at;
    rhs = rhs + at * F.u_old[3 * (size_t)r + comp];
  } else if (q.ts_order == 2) {
    const double at = q.rho * vol / (2.0 * q.dt);
    ap = ap + 3.0 * at;
    rhs = rhs + at * (4.0 * F.u_old[3 * (size_t)r + comp] - F.u_old_old[3 * (size_t)r + comp]);
  }
  if (q.ts_order > 0 && !q.pimple) {
    val[m.diag_slot[r]] = ap;
  } else {
    val[m.diag_slot[r]] = ap / q.alpha_u;
    rhs = rhs + (1.0 - q.alpha_u) / q.alpha_u * ap * uP[comp];
  }
  rhs_out[r] = rhs;
  ap_out[r] = ap;
}

__global__ void k_momentum(DMesh m, DFields F, Ramp rp, MomParams q, const double *psi, double *val, double *rhs_out,
                           double *ap_out) {
  const int r = cell_row(m);
  if (r < 0) return;
  const int comp = q.comp;
  const int k0 = m.cf_ptr[r], k1 = m.cf_ptr[r + 1];
  double ap = 0.0, rhs = 0.0;
  for (int k = k0; k < k1; k++)
    if (m.cf_slot[k] >= 0) val[m.cf_slot[k]] = 0.0;
  // interior faces (:974-1120)
  for (int k = k0; k < k1; k++) {
    const int e = m.cf_ent[k], f = e >> 1, side = e & 1;
    const int o = m.e_oth[k];
    if (o < 0) continue;
    const int a = side == 0 ? r : o, b = side == 0 ? o : r;   // c0 and c1 of the face
    const double Ff = F.flux[f], D = m.e_Dm[k], l = m.e_lf[k];
    const long long sl = m.cf_slot[k];
    const int fl = face_flags(m, f);
    double Dface = D;
    if (fl & FF_BJ) {   // Beavers-Joseph slip at a fluid/porous face (:1031-1059)
      const double Cb = m.Cbj[f], ar = m.area[f];
      double nh[3];
      for (int d = 0; d < 3; d++) nh[d] = m.sf[3 * (size_t)f + d] / ar;
      Dface = Cb + (D - Cb) * (nh[comp] * nh[comp]);
      const double *u0 = &F.u[3 * (size_t)a], *u1 = &F.u[3 * (size_t)b];
      const double rx = (D - Cb) * nh[comp] *
                        ((u0[0] - u1[0]) * nh[0] + (u0[1] - u1[1]) * nh[1] + (u0[2] - u1[2]) * nh[2] -
                         (u0[comp] - u1[comp]) * nh[comp]);
      rhs = side == 0 ? rhs + rx : rhs - rx;
    }
    if (side == 0) {
      ap = ap + dmax(Ff, 0.0) + Dface;
      val[sl] = val[sl] - (Dface + dmax(-Ff, 0.0));
    } else {
      ap = ap + dmax(-Ff, 0.0) + Dface;
      val[sl] = val[sl] - (Dface + dmax(Ff, 0.0));
    }
    if (m.blend_on) {
      const int cu = Ff > 0.0 ? a : b;
      double g[3], rr[3];
      for (int d = 0; d < 3; d++) g[d] = F.gu[9 * (size_t)cu + comp + 3 * d];
      face_offset(m, f, cu, !(Ff > 0.0), rr);
      double dc = psi[cu] * ddot3(g, rr);
      dc = q.blend * Ff * dc;
      rhs = side == 0 ? rhs - dc : rhs + dc;
    }
    if (m.nonorth && !(fl & FF_BJ)) {
      double gf[3];
      for (int d = 0; d < 3; d++)
        gf[d] = l * F.gu[9 * (size_t)a + comp + 3 * d] + (1.0 - l) * F.gu[9 * (size_t)b + comp + 3 * d];
      const double noc = q.nonorth * m.e_muf[k] * ddot3(gf, &m.e_tcor[3 * (size_t)k]);
      rhs = side == 0 ? rhs + noc : rhs - noc;
    }
    const double pf = (fl & FF_KINK) ? kink_pf(m, F.p, F.gp, f, a, b) : l * F.p[a] + (1.0 - l) * F.p[b];
    const double Sc = m.e_sf[3 * (size_t)k + comp];
    rhs = side == 0 ? rhs - pf * Sc : rhs + pf * Sc;
  }
  mom_tail(m, F, rp, q, r, k0, k1, ap, rhs, val, rhs_out, ap_out);
}

// The interior faces of all three components in one pass (the momentum
// matrix's off-diagonals are the same for the three; ap and rhs3 hold each
// component's sums over the interior faces, which k_mom_tail continues).
// Every component's operations are those of k_momentum in the same order,
// so the result is the same bit for bit, provided no interior face term of
// one component reads another component's velocity (no Beavers-Joseph
// faces) and the limiters psi[c * n + r] of all three were computed before
// the first solve (k_limiter3).  The face data, the flux, the neighbours'
// gradients and the matrix slots are read once for the three components.
__global__ void k_mom_interior3(DMesh m, DFields F, MomParams q, const double *psi, double *val, double *rhs3,
                                double *ap1) {
  const int r = cell_row(m);
  if (r < 0) return;
  const int k0 = m.cf_ptr[r], k1 = m.cf_ptr[r + 1];
  double ap = 0.0, rhs[3] = {0.0, 0.0, 0.0};
  for (int k = k0; k < k1; k++)
    if (m.cf_slot[k] >= 0) val[m.cf_slot[k]] = 0.0;
  for (int k = k0; k < k1; k++) {
    const int e = m.cf_ent[k], f = e >> 1, side = e & 1;
    const int o = m.e_oth[k];
    if (o < 0) continue;
    const int a = side == 0 ? r : o, b = side == 0 ? o : r;
    const double Ff = F.flux[f], D = m.e_Dm[k], l = m.e_lf[k];
    const long long sl = m.cf_slot[k];
    const int fl = face_flags(m, f);
    if (side == 0) {
      ap = ap + dmax(Ff, 0.0) + D;
      val[sl] = val[sl] - (D + dmax(-Ff, 0.0));
    } else {
      ap = ap + dmax(-Ff, 0.0) + D;
      val[sl] = val[sl] - (D + dmax(Ff, 0.0));
    }
    if (m.blend_on) {
      const int cu = Ff > 0.0 ? a : b;
      double rr[3];
      face_offset(m, f, cu, !(Ff > 0.0), rr);
      const double *gc = &F.gu[9 * (size_t)cu];
      for (int c = 0; c < 3; c++) {
        double g[3];
        for (int d = 0; d < 3; d++) g[d] = gc[c + 3 * d];
        double dc = psi[(size_t)c * m.nc + cu] * ddot3(g, rr);
        dc = q.blend * Ff * dc;
        rhs[c] = side == 0 ? rhs[c] - dc : rhs[c] + dc;
      }
    }
    if (m.nonorth) {
      const double *ga = &F.gu[9 * (size_t)a], *gb = &F.gu[9 * (size_t)b];
      const double mf = m.e_muf[k];
      const double *tc = &m.e_tcor[3 * (size_t)k];
      for (int c = 0; c < 3; c++) {
        double gf[3];
        for (int d = 0; d < 3; d++) gf[d] = l * ga[c + 3 * d] + (1.0 - l) * gb[c + 3 * d];
        const double noc = q.nonorth * mf * ddot3(gf, tc);
        rhs[c] = side == 0 ? rhs[c] + noc : rhs[c] - noc;
      }
    }
    const double pf = (fl & FF_KINK) ? kink_pf(m, F.p, F.gp, f, a, b) : l * F.p[a] + (1.0 - l) * F.p[b];
    for (int c = 0; c < 3; c++) {
      const double Sc = m.e_sf[3 * (size_t)k + c];
      rhs[c] = side == 0 ? rhs[c] - pf * Sc : rhs[c] + pf * Sc;
    }
  }
  ap1[r] = ap;
  for (int c = 0; c < 3; c++) rhs3[3 * (size_t)r + c] = rhs[c];
}

// component q.comp after k_mom_interior3: boundary faces, sources, diagonal
__global__ void k_mom_tail(DMesh m, DFields F, Ramp rp, MomParams q, const double *rhs3, const double *ap1,
                           double *val, double *rhs_out, double *ap_out) {
  const int r = cell_row(m);
  if (r < 0) return;
  mom_tail(m, F, rp, q, r, m.cf_ptr[r], m.cf_ptr[r + 1], ap1[r], rhs3[3 * (size_t)r + q.comp], val, rhs_out,
           ap_out);
}

// ----------------------------------------------------------------- energy
struct EnParams {
  int ts_order, pimple;
  double rho, cp, alpha_u, dt, blend, nonorth, k_cond;
};

// thermal dispersion conductivity (Bear model, :1781-1799): kdisp(3,n)
__global__ void k_kdisp(DFields F, int n, double rho, double cp) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  double kd[3] = {0.0, 0.0, 0.0};
  const double dl = F.disp_l[r], dt = F.disp_t[r];
  if (dl > 0.0 || dt > 0.0) {
    const double *u = &F.u[3 * (size_t)r];
    const double umag = sqrt(ddot3(u, u));
    if (!(umag < 1.0e-14)) {
      const double umag2 = umag * umag;
      for (int d = 0; d < 3; d++) kd[d] = rho * cp / umag * (dl * (u[d] * u[d]) + dt * (umag2 - u[d] * u[d]));
    }
  }
  for (int d = 0; d < 3; d++) F.kdisp[3 * (size_t)r + d] = kd[d];
}

// n.D.n of a face with the dispersion tensor of cell c: sum (S_d/|S|)^2 kdisp_d
__device__ __forceinline__ double disp_n(const DMesh &m, const DFields &F, int f, int c) {
  double nd[3];
  for (int d = 0; d < 3; d++) {
    const double t = m.sf[3 * (size_t)f + d] / m.area[f];
    nd[d] = t * t;
  }
  return ddot3(nd, &F.kdisp[3 * (size_t)c]);
}

// temperature_assembly (:1748-2037): LTE or the fluid phase of LTNE
__global__ void k_energy(DMesh m, DFields F, EnParams q, const double *psi, double *val, double *rhs_out,
                         double *ap_out) {
  const int r = cell_row(m);
  if (r < 0) return;
  const int k0 = m.cf_ptr[r], k1 = m.cf_ptr[r + 1];
  double ap = 0.0, rhs = 0.0;
  for (int k = k0; k < k1; k++)
    if (m.cf_slot[k] >= 0) val[m.cf_slot[k]] = 0.0;
  for (int k = k0; k < k1; k++) {
    const int e = m.cf_ent[k], f = e >> 1, side = e & 1;
    const int b = m.c1[f];
    if (b < 0) continue;
    const int a = m.c0[f];
    const int fl = face_flags(m, f);
    if (fl & FF_SOLID2) continue;                  // solid-solid: not in the fluid equation
    if (fl & FF_CHT) {                             // fluid side of a solid face (:1812-1834)
      const int o = side == 0 ? b : a;
      if (!solid_cell(F, r)) {
        ap = ap + m.kch[f];
        rhs = rhs + m.kch[f] * F.T_s[o];
      }
      continue;
    }
    const double FT = q.cp * F.flux[f], l = m.lf[f];
    double DT = m.Dt[f], kf = m.kf[f];
    if (m.has_disp) {
      kf = kf + (l * disp_n(m, F, f, a) + (1.0 - l) * disp_n(m, F, f, b));
      DT = m.nonorth ? kf * (m.area[f] * m.area[f]) / m.dvs[f] : kf * m.area[f] / m.dn[f];
    }
    const long long sl = m.cf_slot[k];
    if (side == 0) {
      ap = ap + dmax(FT, 0.0) + DT;
      val[sl] = val[sl] - (DT + dmax(-FT, 0.0));
    } else {
      ap = ap + dmax(-FT, 0.0) + DT;
      val[sl] = val[sl] - (DT + dmax(FT, 0.0));
    }
    if (m.blend_on) {
      const int cu = FT > 0.0 ? a : b;
      double rr[3];
      face_offset(m, f, cu, !(FT > 0.0), rr);
      double dc = psi[cu] * ddot3(&F.gt[3 * (size_t)cu], rr);
      dc = q.blend * FT * dc;
      rhs = side == 0 ? rhs - dc : rhs + dc;
    }
    if (m.nonorth) {
      double gf[3];
      for (int d = 0; d < 3; d++) gf[d] = l * F.gt[3 * (size_t)a + d] + (1.0 - l) * F.gt[3 * (size_t)b + d];
      const double noc = q.nonorth * kf * ddot3(gf, &m.tcor[3 * (size_t)f]);
      rhs = side == 0 ? rhs + noc : rhs - noc;
    }
  }
  const double por = F.porosity[r];
  for (int k = k0; k < k1; k++) {
    const int f = m.cf_ent[k] >> 1;
    if (m.c1[f] >= 0) continue;
    const double FT = q.cp * F.flux[f];
    double DT = m.Dtb[f];
    if (m.has_disp) {
      double kf = m.ltne ? por * q.k_cond : por * q.k_cond + (1.0 - por) * F.k_s[r];
      kf = kf + disp_n(m, F, f, r);
      DT = kf * m.area[f] / m.dw[f];
    }
    const int tm = m.tmode[f];
    if (tm == 1) {
      const double Tf = m.tval[f];
      ap = ap + DT + dmax(FT, 0.0);
      rhs = rhs + DT * Tf - dmin(FT, 0.0) * Tf;
    } else {
      const double qf = tm == 2 ? m.qval[f] : 0.0;
      if (fabs(qf) > 0.0) {
        if (m.ltne) {   // the fluid's conductivity share of the wall flux (:1918-1924)
          const double kft = por * q.k_cond, kst = (1.0 - por) * F.k_s[r];
          rhs = rhs + qf * m.area[f] * kft / (kft + kst);
        } else {
          rhs = rhs + qf * m.area[f];
        }
      }
      const int bt = m.bc[m.fgrp[f] - 1].btype;
      if (bt == MS_BC_POUTLET || bt == MS_BC_OUTFLOW) ap = ap + dmax(FT, 0.0);
    }
  }
  const double vol = m.vol[r];
  if (m.ltne) {    // fluid-solid exchange (:1959-1967)
    const double hasf = F.h_sf[r] * F.a_sf[r];
    if (hasf > 0.0) {
      ap = ap + hasf * vol;
      rhs = rhs + hasf * F.T_s[r] * vol;
    }
  }
  const double cap = m.ltne ? por * q.rho * q.cp : por * q.rho * q.cp + (1.0 - por) * F.rho_s[r] * F.cp_s[r];
  if (q.ts_order == 1) {
    const double at = cap * vol / q.dt;
    ap = ap + at;
    rhs = rhs + at * F.T_old[r];
  } else if (q.ts_order == 2) {
    const double at = cap * vol / (2.0 * q.dt);
    ap = ap + 3.0 * at;
    rhs = rhs + at * (4.0 * F.T_old[r] - F.T_old_old[r]);
  }
  if (q.ts_order > 0 && !q.pimple) {
    val[m.diag_slot[r]] = ap;
  } else {
    val[m.diag_slot[r]] = ap / q.alpha_u;
    rhs = rhs + (1.0 - q.alpha_u) / q.alpha_u * ap * F.T[r];
  }
  if (solid_cell(F, r)) {    // solid cells keep T = T_s (:2024-2033)
    val[m.diag_slot[r]] = 1.0;
    rhs = F.T_s[r];
  }
  rhs_out[r] = rhs;
  ap_out[r] = ap;
}

// solid_temperature_assembly (:2109-2296): the solid phase of LTNE and the
// conduction in solid zones
__global__ void k_solid_energy(DMesh m, DFields F, EnParams q, double *val, double *rhs_out, double *ap_out) {
  const int r = cell_row(m);
  if (r < 0) return;
  const int k0 = m.cf_ptr[r], k1 = m.cf_ptr[r + 1];
  double ap = 0.0, rhs = 0.0;
  for (int k = k0; k < k1; k++)
    if (m.cf_slot[k] >= 0) val[m.cf_slot[k]] = 0.0;
  for (int k = k0; k < k1; k++) {
    const int e = m.cf_ent[k], f = e >> 1, side = e & 1;
    const int b = m.c1[f];
    if (b < 0) continue;
    const int a = m.c0[f];
    if (face_flags(m, f) & FF_CHT) {               // solid side of a fluid face (:2142-2160)
      const int o = side == 0 ? b : a;
      if (solid_cell(F, r)) {
        ap = ap + m.kch[f];
        rhs = rhs + m.kch[f] * F.T[o];
      }
      continue;
    }
    const double Ds = m.Ds[f];
    const long long sl = m.cf_slot[k];
    ap = ap + Ds;
    val[sl] = val[sl] - Ds;
    if (m.nonorth) {
      const double l = m.lf[f];
      double gf[3];
      for (int d = 0; d < 3; d++) gf[d] = l * F.gts[3 * (size_t)a + d] + (1.0 - l) * F.gts[3 * (size_t)b + d];
      const double noc = q.nonorth * m.ksf[f] * ddot3(gf, &m.tcor[3 * (size_t)f]);
      rhs = side == 0 ? rhs + noc : rhs - noc;
    }
  }
  const double por = F.porosity[r];
  for (int k = k0; k < k1; k++) {
    const int f = m.cf_ent[k] >> 1;
    if (m.c1[f] >= 0) continue;
    const int tm = m.tmode_s[f];
    if (tm == 1) {
      ap = ap + m.Dsb[f];
      rhs = rhs + m.Dsb[f] * m.tval_s[f];
    } else {
      const double qf = tm == 2 ? m.qval[f] : 0.0;
      if (fabs(qf) > 0.0) {
        const double ksf = (1.0 - por) * F.k_s[r], kft = por * q.k_cond;
        rhs = rhs + qf * m.area[f] * ksf / (kft + ksf);
      }
    }
  }
  const double vol = m.vol[r];
  const double hasf = F.h_sf[r] * F.a_sf[r];
  if (hasf > 0.0) {
    ap = ap + hasf * vol;
    rhs = rhs + hasf * F.T[r] * vol;
  }
  if (F.q_src[r] != 0.0) rhs = rhs + F.q_src[r] * vol;
  if (q.ts_order == 1) {
    const double at = (1.0 - por) * F.rho_s[r] * F.cp_s[r] * vol / q.dt;
    ap = ap + at;
    rhs = rhs + at * F.T_s_old[r];
  } else if (q.ts_order == 2) {
    const double at = (1.0 - por) * F.rho_s[r] * F.cp_s[r] * vol / (2.0 * q.dt);
    ap = ap + 3.0 * at;
    rhs = rhs + at * (4.0 * F.T_s_old[r] - F.T_s_old_old[r]);
  }
  if (q.ts_order > 0 && !q.pimple) {
    val[m.diag_slot[r]] = ap;
  } else {
    val[m.diag_slot[r]] = ap / q.alpha_u;
    rhs = rhs + (1.0 - q.alpha_u) / q.alpha_u * ap * F.T_s[r];
  }
  if (m.has_solid && !solid_cell(F, r)) {   // pure fluid cells keep T_s (:2280-2292)
    const double *K = &F.perm_dir[3 * (size_t)r];
    const bool porous = K[0] > 0.0 || K[1] > 0.0 || K[2] > 0.0 || F.inertial[r] > 0.0;
    if (!porous) {
      val[m.diag_slot[r]] = 1.0;
      rhs = F.T_s[r];
    }
  }
  rhs_out[r] = rhs;
  ap_out[r] = ap;
}

// -------------------------------------------------------------- Rhie-Chow
__global__ void k_rhiechow(DMesh m, DFields F, Ramp rp, double rho) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.nf) return;
  const int a = m.c0[f], b = m.c1[f];
  const double *sf = &m.sf[3 * (size_t)f];
  if (b < 0) {
    if (m.fgrp[f] <= 0) return;   // shadow of a joined periodic face
    double uf[3];
    bc_face_vel(m, f, &F.u[3 * (size_t)a], uf, rp);
    F.flux[f] = rho * ddot3(uf, sf);
    return;
  }
  const double l = m.lf[f];
  double ufl[3], gpl[3];
  for (int d = 0; d < 3; d++) {
    ufl[d] = l * F.u[3 * (size_t)a + d] + (1.0 - l) * F.u[3 * (size_t)b + d];
    gpl[d] = l * F.gp[3 * (size_t)a + d] + (1.0 - l) * F.gp[3 * (size_t)b + d];
  }
  const double dn = m.dn[f];
  const double pfn = (F.p[b] - F.p[a]) / dn;
  const double dbf = l * m.vol[a] / F.apc[a] + (1.0 - l) * m.vol[b] / F.apc[b];
  F.flux[f] = rho * (ddot3(ufl, sf) + dbf * (ddot3(gpl, sf) - pfn * m.dvs[f] / dn));
}

// ----------------------------------------------------------- outflow rescale
__global__ void k_outflow_sums(DMesh m, const double *flux, double *part) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (f < m.nf && m.c1[f] < 0 && m.fgrp[f] > 0) {
    int bt = m.bc[m.fgrp[f] - 1].btype;
    if (bt == MS_BC_OUTFLOW) {
      v[1] = flux[f];
      v[2] = m.area[f];
    } else if (bt != MS_BC_POUTLET && bt != MS_BC_FARFIELD) {
      v[0] = -flux[f];
    }
  }
  block_partials<OpSum, 3>(v, part);
}

struct FinStore {      // copy the reduced values into out[0..nv)
  double *out;
  int nv;
  __device__ bool skip() const { return false; }
  __device__ void operator()(const double *a) const {
    for (int j = 0; j < nv; j++) out[j] = a[j];
  }
};

__global__ void k_outflow_scale(DMesh m, double *flux, const double *sums) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.nf || m.c1[f] >= 0 || m.fgrp[f] <= 0) return;
  if (m.bc[m.fgrp[f] - 1].btype != MS_BC_OUTFLOW) return;
  const double m_req = sums[0], m_out = sums[1], a_out = sums[2];
  if (a_out <= 0.0 || m_req <= 0.0) return;
  if (m_out > 1.0e-2 * m_req) flux[f] = (m_req / m_out) * flux[f];
  else flux[f] = m_req * m.area[f] / a_out;
}

// ------------------------------------------------------------------- PPE
__global__ void k_ppe_af(DMesh m, const double *apc, double rho, double *af) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.nf) return;
  const int a = m.c0[f], b = m.c1[f];
  if (b >= 0) {
    const double l = m.lf[f];
    const double dbf = l * m.vol[a] / apc[a] + (1.0 - l) * m.vol[b] / apc[b];
    af[f] = rho * dbf * m.area[f] / m.dn[f];
  } else if (p_dirichlet(m, f)) {
    const double dbf = m.vol[a] / apc[a];
    af[f] = rho * dbf * m.area[f] / m.dw[f];
  } else {
    af[f] = 0.0;
  }
}

__global__ void k_ppe(DMesh m, const double *flux, const double *af, double *val, double *rhs_out) {
  const int r = cell_row(m);
  if (r < 0) return;
  const int k0 = m.cf_ptr[r], k1 = m.cf_ptr[r + 1];
  for (int k = k0; k < k1; k++)
    if (m.cf_slot[k] >= 0) val[m.cf_slot[k]] = 0.0;
  double diag = 0.0, rhs = 0.0;
  for (int k = k0; k < k1; k++) {
    const int e = m.cf_ent[k], f = e >> 1, side = e & 1;
    if (m.e_oth[k] < 0) {
      if (p_dirichlet_ppe(m, f)) diag = diag + af[f];
      rhs = rhs - flux[f];
      continue;
    }
    diag = diag + af[f];
    const long long sl = m.cf_slot[k];
    val[sl] = val[sl] - af[f];
    rhs = side == 0 ? rhs - flux[f] : rhs + flux[f];
  }
  if (!m.has_pdir) {           // pin cell 1 (:1629-1642)
    if (r == m.pin_cell) {
      for (int k = k0; k < k1; k++)
        if (m.cf_slot[k] >= 0) val[m.cf_slot[k]] = 0.0;
      diag = 1.0;
      rhs = 0.0;
    } else {
      for (int k = k0; k < k1; k++) {
        const int o = m.e_oth[k];
        if (o < 0) continue;
        if (o == m.pin_cell) val[m.cf_slot[k]] = 0.0;
      }
    }
  }
  val[m.diag_slot[r]] = diag;
  rhs_out[r] = rhs;
}

// ------------------------------------------------------------- correction
__global__ void k_correct_cells(DMesh m, DFields F, const double *pp, double *dpp_out, double palpha) {
  const int r = cell_row(m);
  if (r < 0) return;
  double g[3] = {0.0, 0.0, 0.0};
  for (int k = m.cf_ptr[r]; k < m.cf_ptr[r + 1]; k++) {
    const int e = m.cf_ent[k], f = e >> 1, side = e & 1;
    const int o = m.e_oth[k];
    double phi;
    if (o >= 0) {
      const double l = m.e_lf[k];
      const int a = side == 0 ? r : o, b = side == 0 ? o : r;   // c0 and c1 of the face
      phi = l * pp[a] + (1.0 - l) * pp[b];
    } else if (p_dirichlet(m, f)) {
      phi = 0.0;
    } else {
      phi = pp[r];
    }
    const double *sf = &m.e_sf[3 * (size_t)k];
    if (side == 0)
      for (int d = 0; d < 3; d++) g[d] = g[d] + phi * sf[d];
    else
      for (int d = 0; d < 3; d++) g[d] = g[d] - phi * sf[d];
  }
  const double v = m.vol[r];
  for (int d = 0; d < 3; d++) g[d] = g[d] / v;
  const double s = v / F.apc[r];
  for (int d = 0; d < 3; d++) {
    F.u[3 * (size_t)r + d] = F.u[3 * (size_t)r + d] - s * g[d];
    dpp_out[3 * (size_t)r + d] = g[d];
  }
  F.p[r] = palpha < 0.0 ? F.p[r] + pp[r] : F.p[r] + palpha * pp[r];
  if (solid_cell(F, r))   // zero_solid_velocity (:2064-2076)
    for (int d = 0; d < 3; d++) F.u[3 * (size_t)r + d] = F.u_old[3 * (size_t)r + d] = 0.0;
}

__global__ void k_zero_solid(DFields F, int n) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n && solid_cell(F, r))
    for (int d = 0; d < 3; d++) F.u[3 * (size_t)r + d] = F.u_old[3 * (size_t)r + d] = 0.0;
}

__global__ void k_correct_flux(DMesh m, const double *pp, const double *af, double *flux) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.nf) return;
  const int a = m.c0[f], b = m.c1[f];
  if (b < 0) {
    if (p_dirichlet(m, f)) flux[f] = flux[f] + af[f] * pp[a];
    return;
  }
  flux[f] = flux[f] - af[f] * (pp[b] - pp[a]);
}

// ------------------------------------------------------------------ small
__global__ void k_comp_get(int n, const double *u, int comp, double *x) {
  int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) x[r] = u[3 * (size_t)r + comp];
}
__global__ void k_comp_set(int n, double *u, int comp, const double *x) {
  int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) u[3 * (size_t)r + comp] = x[r];
}
__global__ void k_apc(int n, const double *ap, double alpha, double *apc) {
  int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) apc[r] = alpha > 0.0 ? ap[r] / alpha : ap[r];
}
__global__ void k_fill(long long n, double *x, double v) {
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = v;
}

// y = x / d (the relaxed momentum diagonal apc = ap / alpha_u)
__global__ void k_div(int n, const double *x, double d, double *y) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = x[i] / d;
}

// ------------------------------------------------------------ diagnostics
// out: [0] max|imbalance| [1] max|u - uref| [2] max|T - Tref| [3] max norm2(u) [4] non-finite
__global__ void k_diagnostics(DMesh m, DFields F, const double *uref, const double *Tref, double *part) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (r < m.nc) {
    double imb = 0.0;
    for (int k = m.cf_ptr[r]; k < m.cf_ptr[r + 1]; k++) {
      const int e = m.cf_ent[k], f = e >> 1;
      imb = (e & 1) == 0 ? imb + F.flux[f] : imb - F.flux[f];
    }
    v[0] = fabs(imb);
    const double *u = &F.u[3 * (size_t)r];
    for (int d = 0; d < 3; d++) v[1] = dmax(v[1], fabs(u[d] - uref[3 * (size_t)r + d]));
    v[2] = fabs(F.T[r] - Tref[r]);
    v[3] = dnorm3(u);
    bool bad = !isfinite(u[0]) || !isfinite(u[1]) || !isfinite(u[2]) || !isfinite(F.p[r]) || !isfinite(F.T[r]) ||
               !isfinite(F.T_s[r]);   // nan_scan covers T_s (mod_uns_simple.f90:2402-2404)
    v[4] = bad ? 1.0 : 0.0;
  }
  block_partials<OpMax, 5>(v, part);
}

// per wall face: pressure force (p - p_inf) S, and with mu > 0 the wall shear
// tau = mu u_t / dn along t = n x z_hat times the face area
__global__ void k_wall_forces(DMesh m, DFields F, const int *wf, int nw, double mu, double p_inf, double *part) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (t < nw) {
    const int f = wf[t], c = m.c0[f];
    const double *sf = &m.sf[3 * (size_t)f], *xf = &m.xf[3 * (size_t)f];
    const double pg = F.p[c] - p_inf;
    v[0] = pg * sf[0];
    v[1] = pg * sf[1];
    v[2] = pg * sf[2];
    if (mu > 0.0) {
      const double nx = sf[0] / dnorm3(sf), ny = sf[1] / dnorm3(sf);
      const double tx = ny, ty = -nx, tz = 0.0;
      const double d[3] = {xf[0] - m.xc[3 * (size_t)c], xf[1] - m.xc[3 * (size_t)c + 1],
                           xf[2] - m.xc[3 * (size_t)c + 2]};
      const double dn = dnorm3(d);
      const double *u = &F.u[3 * (size_t)c];
      const double ut = u[0] * tx + u[1] * ty + u[2] * tz;
      const double tau = mu * ut / dn;
      v[3] = tau * tx * m.area[f];
      v[4] = tau * ty * m.area[f];
      v[5] = tau * tz * m.area[f];
    }
  }
  block_partials<OpSum, 6>(v, part);
}

// (p, u) of the owner cell of every wall face
__global__ void k_wall_samples(DMesh m, DFields F, const int *wf, int nw, double *out) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nw) return;
  const int c = m.c0[wf[t]];
  out[4 * (size_t)t] = F.p[c];
  out[4 * (size_t)t + 1] = F.u[3 * (size_t)c];
  out[4 * (size_t)t + 2] = F.u[3 * (size_t)c + 1];
  out[4 * (size_t)t + 3] = F.u[3 * (size_t)c + 2];
}

// farthest reverse-flow cell behind x0: max of x0 and xc over cells with u_x < 0, xc > x0
__global__ void k_wake_end(DMesh m, DFields F, double x0, double *part) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  double v[1] = {x0};
  if (r < m.nc && F.u[3 * (size_t)r] < 0.0 && m.xc[3 * (size_t)r] > x0) v[0] = dmax(x0, m.xc[3 * (size_t)r]);
  block_partials<OpMax, 1>(v, part);
}

}  // namespace

// ======================================================================
// host drivers
// ======================================================================
static dim3 gc(int n) { return dim3(nblocks(n)); }
// grid of the cell-wise assembly kernels (cell_row)
static dim3 gcell(const System &s) {
  if (s.P->color_ptr.size() == 3) {
    const int n0 = s.P->color_ptr[1];
    return dim3(nblocks(std::max(n0, s.nc - n0), kBlock / 2));
  }
  return gc(s.nc);
}

// The preconditioner of a solve: natural and multicolour selectors give the
// level-scheduled factorization in the system's numbering (IC(0) under PCG,
// ILU(0) under BiCGSTAB); with the colour-blocked numbering its levels are
// the colour classes.
static PcKind pc_of(int sel, int solver) {
  if (sel == MS_PC_JACOBI) return PC_JACOBI;
  if (sel == MS_PC_DILU) return PC_DILU;
  return solver == 0 ? PC_NATURAL_IC0 : PC_NATURAL_ILU0;
}

struct Driver {
  System &s;
  Ramp rp{1.0, 1.0};
  explicit Driver(System &sys) : s(sys) { rp = Ramp{s.inlet_ramp, s.pval_ramp}; }

  void gradients() {
    if (s.has_kink)   // kink faces reconstruct with the gradient of the previous pass
      MS_HIP(hipMemcpyAsync(s.gp_prev.p, s.gp.p, 3 * (size_t)s.nc * sizeof(double), hipMemcpyDeviceToDevice,
                            stream()));
    hipLaunchKernelGGL(k_gradients, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), rp,
                       (int)s.energy_on());
    MS_CHECK_LAUNCH();
  }
  void rhiechow() {
    hipLaunchKernelGGL(k_rhiechow, gc(s.nf), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), rp, s.ctrl.rho);
    MS_CHECK_LAUNCH();
  }
  void outflow() {
    if (!s.has_outflow) return;
    hipLaunchKernelGGL(k_outflow_sums, gc(s.nf), dim3(kBlock), 0, stream(), s.dmesh(), s.flux.p, s.part.p);
    hipLaunchKernelGGL((k_finish<OpSum, 3, double, FinStore>), dim3(1), dim3(kFinThreads), 0, stream(), nblocks(s.nf),
                       s.part.p, FinStore{s.tmpc.p, 3});
    hipLaunchKernelGGL(k_outflow_scale, gc(s.nf), dim3(kBlock), 0, stream(), s.dmesh(), s.flux.p, s.tmpc.p);
    MS_CHECK_LAUNCH();
  }
  MomParams mom_params(int comp) const {
    const ms_ctrl_t &c = s.ctrl;
    MomParams q{};
    q.comp = comp;
    q.ts_order = s.ts_order;
    q.pimple = c.pimple;
    q.boussinesq = c.boussinesq;
    q.body_on = c.body_force[0] != 0.0 || c.body_force[1] != 0.0 || c.body_force[2] != 0.0;
    q.rho = c.rho;
    q.mu = c.mu;
    q.alpha_u = c.alpha_u;
    q.dt = c.dt;
    q.beta = c.beta;
    q.tref = c.tref;
    q.blend = c.conv_blend;
    q.nonorth = c.nonorth_corr;
    q.grav = c.gravity[comp];
    q.body = c.body_force[comp];
    return q;
  }
  void momentum(int comp) {     // comp 0..2
    if (s.ctrl.conv_blend > 0.0)
      hipLaunchKernelGGL(k_limiter, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), rp, comp, s.psi.p);
    hipLaunchKernelGGL(k_momentum, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), rp, mom_params(comp),
                       s.psi.p, s.val.p, s.rhs.p, s.ap.p);
    MS_CHECK_LAUNCH();
  }
  // The three components' interior faces in one pass (k_mom_interior3) is
  // exact when no interior or limiter term of one component reads another
  // component's velocity: no Beavers-Joseph faces, and with the limiter on
  // no symmetry or slip-wall group (their face velocity mixes components).
  bool momentum_fused() const {
    if (!momentum_fused_enabled() || s.has_bj) return false;
    if (s.ctrl.conv_blend > 0.0)
      for (const ms_bcgroup_t &g : s.groups)
        if (g.btype == MS_BC_SYMMETRY || g.btype == MS_BC_SLIPWALL) return false;
    return true;
  }
  void momentum_interior3() {
    if (s.mom_rhs3.n < 3 * (size_t)s.nc) s.mom_rhs3.alloc(3 * (size_t)s.nc);
    if (s.mom_ap1.n < (size_t)s.nc) s.mom_ap1.alloc(s.nc);
    if (s.ctrl.conv_blend > 0.0)
      hipLaunchKernelGGL(k_limiter3, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), rp, s.psi.p);
    hipLaunchKernelGGL(k_mom_interior3, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), mom_params(0),
                       s.psi.p, s.val.p, s.mom_rhs3.p, s.mom_ap1.p);
    MS_CHECK_LAUNCH();
  }
  void momentum_tail(int comp) {
    hipLaunchKernelGGL(k_mom_tail, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), rp, mom_params(comp),
                       s.mom_rhs3.p, s.mom_ap1.p, s.val.p, s.rhs.p, s.ap.p);
    MS_CHECK_LAUNCH();
  }
  void energy() {
    const ms_ctrl_t &c = s.ctrl;
    if (c.conv_blend > 0.0)
      hipLaunchKernelGGL(k_limiter, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), rp, 3, s.psi.p);
    EnParams q{s.ts_order, c.pimple, c.rho, c.cp, c.alpha_u, c.dt, c.conv_blend, c.nonorth_corr, c.k_cond};
    if (s.has_disp)
      hipLaunchKernelGGL(k_kdisp, gc(s.nc), dim3(kBlock), 0, stream(), s.dfields(), s.nc, c.rho, c.cp);
    hipLaunchKernelGGL(k_energy, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), q, s.psi.p, s.val.p,
                       s.rhs.p, s.ap.p);
    MS_CHECK_LAUNCH();
  }
  void solid_energy() {
    const ms_ctrl_t &c = s.ctrl;
    EnParams q{s.ts_order, c.pimple, c.rho, c.cp, c.alpha_u, c.dt, c.conv_blend, c.nonorth_corr, c.k_cond};
    hipLaunchKernelGGL(k_solid_energy, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), q, s.val.p,
                       s.rhs.p, s.ap.p);
    MS_CHECK_LAUNCH();
  }
  void zero_solid() {
    if (!s.has_solid) return;
    hipLaunchKernelGGL(k_zero_solid, gc(s.nc), dim3(kBlock), 0, stream(), s.dfields(), s.nc);
    MS_CHECK_LAUNCH();
  }
  // solid-phase temperature solve (BiCGSTAB with the momentum preconditioner)
  int solid_solve(int &itl_max) {
    solid_energy();
    SolveResult R = solve(false, s.T_s.p);
    itl_max = std::max(itl_max, R.it);
    return R.ierr;
  }
  void ppe() {
    hipLaunchKernelGGL(k_ppe_af, gc(s.nf), dim3(kBlock), 0, stream(), s.dmesh(), s.apc.p, s.ctrl.rho, s.af.p);
    hipLaunchKernelGGL(k_ppe, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.flux.p, s.af.p, s.val.p, s.rhs.p);
    MS_CHECK_LAUNCH();
  }
  void correct(const double *pp) {
    double palpha = s.ts_order > 0 ? -1.0 : s.ctrl.alpha_p;
    hipLaunchKernelGGL(k_correct_cells, gcell(s), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), pp, s.dpp.p,
                       palpha);
    hipLaunchKernelGGL(k_correct_flux, gc(s.nf), dim3(kBlock), 0, stream(), s.dmesh(), pp, s.af.p, s.flux.p);
    MS_CHECK_LAUNCH();
  }
  // reuse: the matrix equals the one of the previous solve (a later p'
  // corrector with unchanged apc), so the solver keeps its set-up
  SolveResult solve(bool ppe_eq, double *x, bool reuse = false) {
    const ms_ctrl_t &c = s.ctrl;
    int solver = ppe_eq ? c.gpu_ppe_solver : 1;
    PcKind pc = pc_of(ppe_eq ? c.gpu_ppe_precond : c.gpu_mom_precond, solver);
    const bool coarse = ppe_eq && s.P->coarse.k > 0;   // two-level preconditioner for p'
    SolveResult R;
    if (ppe_eq && solver == 2) {   // red-black reduced system (ms_schur.hip)
      R = schur_solve(s, s.val.p, s.rhs.p, x, c.lin_tol, c.lin_max, c.gpu_precision == MS_MIXED, c.gpu_inner_tol,
                      reuse);
    } else if (c.gpu_precision == MS_MIXED) {
      R = s.kry->mixed(s.val.p, s.rhs.p, x, c.lin_tol, c.lin_max, pc, solver, c.gpu_inner_tol, coarse, reuse);
    } else {
      R = solver == 0 ? s.kry->pcg(s.val.p, s.rhs.p, x, c.lin_tol, c.lin_max, pc, coarse, reuse)
                      : s.kry->bicgstab(s.val.p, s.rhs.p, x, c.lin_tol, c.lin_max, pc, coarse);
      R.it64 = R.it;
    }
    if (ppe_eq) {
      s.st_ppe_solves++;
      s.st_it32 += R.it32;
      s.st_it64 += R.it64;
      s.st_fallback += R.stalled;
    } else {
      s.st_mom_solves++;
      s.st_mom_its += R.it;
    }
    return R;
  }
  // momentum predictor over the 3 components; returns failing comp (1..3) or 0
  int momentum_solves(int &itl_max, double &ressum) {
    const bool fused = momentum_fused();
    if (fused) momentum_interior3();
    for (int comp = 0; comp < 3; comp++) {
      if (fused) momentum_tail(comp);
      else momentum(comp);
      hipLaunchKernelGGL(k_comp_get, gc(s.nc), dim3(kBlock), 0, stream(), s.nc, s.u.p, comp, s.ucomp.p);
      SolveResult R = solve(false, s.ucomp.p);
      hipLaunchKernelGGL(k_comp_set, gc(s.nc), dim3(kBlock), 0, stream(), s.nc, s.u.p, comp, s.ucomp.p);
      MS_CHECK_LAUNCH();
      if (R.ierr == 2) return comp + 1;
      itl_max = std::max(itl_max, R.it);
      ressum = ressum + R.res;
    }
    return 0;
  }
  // diagnostics: returns out[5]
  void diagnostics(const double *uref, const double *Tref, double *out) {
    hipLaunchKernelGGL(k_diagnostics, gc(s.nc), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), uref, Tref,
                       s.part.p);
    hipLaunchKernelGGL((k_finish<OpMax, 5, double, FinStore>), dim3(1), dim3(kFinThreads), 0, stream(), nblocks(s.nc),
                       s.part.p, FinStore{s.tmpc.p, 5});
    MS_CHECK_LAUNCH();
    MS_HIP(hipMemcpyAsync(out, s.tmpc.p, 5 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MS_HIP(hipStreamSynchronize(stream()));
  }
};

static void dcopy(double *dst, const double *src, size_t n) {
  MS_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
}

static void fill(double *x, size_t n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(nblocks((long long)n)), dim3(kBlock), 0, stream(), (long long)n, x, v);
  MS_CHECK_LAUNCH();
}

static void push_history(System &s, double a, double b, double c, double d, double e, double f, double g, double h) {
  double row[8] = {a, b, c, d, e, f, g, h};
  s.history.insert(s.history.end(), row, row + 8);
}

// simple_run (mod_uns_simple.f90:88-288) with the solver norms of the
// partitioned path (sqrt of the global dot product).
int run_simple(System &s, int nsteps, int *it_done) {
  s.ppe_serial = true;   // simple_run assembles the serial ppe_assembly
  s.refresh_face_constants();
  Driver D(s);
  const ms_ctrl_t &c = s.ctrl;
  const size_t n3 = 3 * (size_t)s.nc;
  s.history.clear();
  fill(s.apc.p, s.nc, 1.0);
  s.ts_order = 0;
  D.gradients();
  int it_max = c.outer_max;
  if (nsteps > 0 && nsteps < it_max) it_max = nsteps;
  int ier = 0, it;
  for (it = 1; it <= it_max; it++) {
    dcopy(s.u_old.p, s.u.p, n3);      // uo (u_old is unused by the steady assembly)
    int itl_max = 0;
    double ressum = 0.0;
    if (c.inlet_ramp > 1) {
      double f = (double)it / (double)c.inlet_ramp;
      f = std::max(0.0, std::min(1.0, f));
      s.inlet_ramp = s.pval_ramp = f;
      D.rp = Ramp{f, f};
    }
    if (D.momentum_solves(itl_max, ressum)) { ier = 20; break; }
    hipLaunchKernelGGL(k_apc, gc(s.nc), dim3(kBlock), 0, stream(), s.nc, s.ap.p, c.alpha_u, s.apc.p);
    D.zero_solid();
    D.gradients();
    D.rhiechow();
    D.outflow();
    D.ppe();
    fill(s.pp.p, s.nc, 0.0);
    SolveResult R = D.solve(true, s.pp.p);
    if (R.ierr == 2) { ier = 21; break; }
    itl_max = std::max(itl_max, R.it);
    D.correct(s.pp.p);
    dcopy(s.T_old.p, s.T.p, s.nc);    // Tprev
    if (s.energy_on()) {
      D.energy();
      SolveResult RT = D.solve(false, s.T.p);
      if (RT.ierr == 2) { ier = 22; break; }
      itl_max = std::max(itl_max, RT.it);
      // solid-phase temperature: LTNE or solid zones (mod_uns_simple.f90:214-222)
      if (s.solid_phase() && D.solid_solve(itl_max) == 2) { ier = 23; break; }
    }
    double dg[5];
    D.diagnostics(s.u_old.p, s.T_old.p, dg);
    if (dg[4] != 0.0) { ier = 60; break; }
    const double imbal = dg[0] / s.flux_ref, du = dg[1], dT = dg[2];
    push_history(s, imbal, du, dT, 0.0, itl_max, ressum / 3.0, R.it, R.res);
    if (imbal < c.outer_tol && du < c.outer_tol && dT < c.outer_tol) break;
  }
  if (it_done) *it_done = it;
  s.inlet_ramp = s.pval_ramp = 1.0;
  MS_HIP(hipStreamSynchronize(stream()));
  return ier;
}

// piso_run_mpi at one partition (mod_uns_simple_mpi.f90:298-571)
int run_piso(System &s, int nsteps, bool resume) {
  const ms_ctrl_t &c = s.ctrl;
  if (!c.transient) return 30;
  if (c.dt <= 0.0) return 31;
  if (resume && s.steps_done <= 0) fail(MS_EINVAL, "ms_piso_continue: no earlier ms_piso_run on this system");
  s.ppe_serial = false;  // piso_run_mpi: ppe_assembly_mpi
  s.refresh_face_constants();
  Driver D(s);
  const size_t n3 = 3 * (size_t)s.nc;
  int n_steps = std::max(1, c.n_time_max);
  if (nsteps > 0) n_steps = nsteps;
  s.history.clear();
  // a continued run keeps apc, the gradients and the fluxes the last step
  // left, as the reference's loop does from one step to the next
  const int it0 = resume ? s.steps_done : 0;
  if (!resume) {
    s.steps_done = 0;
    fill(s.apc.p, s.nc, 1.0);
    D.gradients();
    D.rhiechow();
  }
  // CFL needs min over cells of vol^(1/3) (host pow, as the reference)
  const double minv13 = s.min_vol13();
  int ier = 0;
  for (int it = it0 + 1; it <= it0 + n_steps; it++) {
    dcopy(s.u_old_old.p, s.u_old.p, n3);
    dcopy(s.u_old.p, s.u.p, n3);
    dcopy(s.T_old_old.p, s.T_old.p, s.nc);
    dcopy(s.T_old.p, s.T.p, s.nc);
    dcopy(s.T_s_old_old.p, s.T_s_old.p, s.nc);
    dcopy(s.T_s_old.p, s.T_s.p, s.nc);
    s.ts_order = (it == 1 || c.time_scheme == 1) ? 1 : 2;
    int itl_max = 0, ppe_it = 0;
    double ressum = 0.0, ppe_res = 0.0;
    if (D.momentum_solves(itl_max, ressum)) { ier = 33; break; }
    dcopy(s.apc.p, s.ap.p, s.nc);
    D.gradients();
    D.rhiechow();
    D.outflow();
    bool bad = false;
    for (int ic = 1; ic <= c.n_correct; ic++) {
      D.ppe();
      fill(s.pp.p, s.nc, 0.0);
      SolveResult R = D.solve(true, s.pp.p, ic > 1);
      if (R.ierr == 2) { bad = true; break; }
      itl_max = std::max(itl_max, R.it);
      ppe_it += R.it;
      ppe_res = std::max(ppe_res, R.res);   // largest over the correctors
      D.correct(s.pp.p);
    }
    if (bad) { ier = 34; break; }
    if (s.energy_on()) {
      D.energy();
      SolveResult RT = D.solve(false, s.T.p);
      if (RT.ierr == 2) { ier = 35; break; }
      itl_max = std::max(itl_max, RT.it);
      // solid-phase temperature: LTNE (mod_uns_simple_mpi.f90:483-495)
      if (c.thermal_model != 0 && D.solid_solve(itl_max) == 2) { ier = 36; break; }
    }
    double dg[5];
    D.diagnostics(s.u_old.p, s.T_old.p, dg);
    const double imbal = dg[0] / s.flux_ref, du = dg[1] / c.dt, umax = dg[3];
    const double cfl = umax * c.dt / minv13;
    if (dg[4] != 0.0) { ier = 60; break; }
    push_history(s, imbal, du, umax, cfl, itl_max, ressum / 3.0, ppe_it, ppe_res);
    s.steps_done = it;
  }
  MS_HIP(hipStreamSynchronize(stream()));
  return ier;
}

// pimple_run_mpi at one partition (mod_uns_simple_mpi.f90:580-850): outer
// iterations per time step, apc = ap / alpha_u (:706), corrections inside
// each outer iteration, energy once per step (:751-760), ts_order kept at 2
// on the first step when the state on entry is a BDF2 restart (:664-682).
int run_pimple(System &s, int nsteps, bool resume) {
  const ms_ctrl_t &c = s.ctrl;
  if (!c.transient) return 50;
  if (c.dt <= 0.0) return 51;
  if (resume && s.steps_done <= 0) fail(MS_EINVAL, "ms_pimple_continue: no earlier ms_pimple_run on this system");
  s.ppe_serial = false;  // pimple_run_mpi: ppe_assembly_mpi
  s.refresh_face_constants();
  Driver D(s);
  const size_t n3 = 3 * (size_t)s.nc;
  int n_steps = std::max(1, c.n_time_max);
  if (nsteps > 0) n_steps = nsteps;
  const int n_outer = std::max(1, c.n_outer_iter);
  s.history.clear();
  const int it0 = resume ? s.steps_done : 0;
  if (!resume) {
    s.steps_done = 0;
    s.ts_order_in = s.ts_order;
    fill(s.apc.p, s.nc, 1.0);
    D.gradients();
    D.rhiechow();
  }
  const int ts_order_in = s.ts_order_in;
  const double minv13 = s.min_vol13();
  int ier = 0;
  for (int it = it0 + 1; it <= it0 + n_steps; it++) {
    dcopy(s.u_old_old.p, s.u_old.p, n3);
    dcopy(s.u_old.p, s.u.p, n3);
    dcopy(s.T_old_old.p, s.T_old.p, s.nc);
    dcopy(s.T_old.p, s.T.p, s.nc);
    dcopy(s.T_s_old_old.p, s.T_s_old.p, s.nc);
    dcopy(s.T_s_old.p, s.T_s.p, s.nc);
    if (c.time_scheme == 1) s.ts_order = 1;
    else if (it == 1 && ts_order_in < 2) s.ts_order = 1;
    else s.ts_order = 2;
    int itl_max = 0, ppe_it = 0;
    double ressum = 0.0, ppe_res = 0.0;
    bool brk = false, bad = false;
    for (int io = 1; io <= n_outer && !brk && !bad; io++) {
      if (D.momentum_solves(itl_max, ressum)) { brk = true; break; }
      hipLaunchKernelGGL(k_div, gc(s.nc), dim3(kBlock), 0, stream(), s.nc, s.ap.p, c.alpha_u, s.apc.p);
      MS_CHECK_LAUNCH();
      D.gradients();
      D.rhiechow();
      D.outflow();
      for (int ic = 1; ic <= c.n_correct; ic++) {
        D.ppe();
        fill(s.pp.p, s.nc, 0.0);
        SolveResult R = D.solve(true, s.pp.p, ic > 1);
        if (R.ierr == 2) { bad = true; break; }
        itl_max = std::max(itl_max, R.it);
        ppe_it += R.it;
        ppe_res = std::max(ppe_res, R.res);   // largest over the correctors
        D.correct(s.pp.p);
      }
    }
    if (brk) { ier = 53; break; }
    if (bad) { ier = 54; break; }
    if (s.energy_on()) {
      D.energy();
      SolveResult RT = D.solve(false, s.T.p);
      if (RT.ierr == 2) { ier = 55; break; }
      itl_max = std::max(itl_max, RT.it);
      // solid-phase temperature: LTNE (mod_uns_simple_mpi.f90:765-777)
      if (c.thermal_model != 0 && D.solid_solve(itl_max) == 2) { ier = 56; break; }
    }
    double dg[5];
    D.diagnostics(s.u_old.p, s.T_old.p, dg);
    const double imbal = dg[0] / s.flux_ref, du = dg[1] / c.dt, umax = dg[3];
    const double cfl = umax * c.dt / minv13;
    if (dg[4] != 0.0) { ier = 60; break; }
    push_history(s, imbal, du, umax, cfl, itl_max, ressum / (3.0 * n_outer), ppe_it, ppe_res);
    s.steps_done = it;
  }
  MS_HIP(hipStreamSynchronize(stream()));
  return ier;
}

// ---- per-call level -------------------------------------------------------
void call_gradients(System &s) {
  s.refresh_face_constants();
  Driver(s).gradients();
}
void call_momentum(System &s, int comp) {
  s.refresh_face_constants();
  Driver(s).momentum(comp);
}
void call_energy(System &s) {
  s.refresh_face_constants();
  Driver(s).energy();
}
void call_solid_energy(System &s) {
  if (!s.solid_consts) {   // the per-call routine also works without an active solid phase
    s.solid_consts = true;
    s.props_dirty = true;
  }
  s.refresh_face_constants();
  Driver(s).solid_energy();
}
void call_rhiechow(System &s) {
  s.refresh_face_constants();
  Driver(s).rhiechow();
}
void call_outflow(System &s) { Driver(s).outflow(); }
void call_ppe(System &s) {
  s.ppe_serial = true;   // the per-call routine is the serial ppe_assembly
  s.refresh_face_constants();
  Driver(s).ppe();
}
void call_correct(System &s, const double *d_pp) {
  s.refresh_face_constants();
  Driver D(s);
  hipLaunchKernelGGL(k_ppe_af, gc(s.nf), dim3(kBlock), 0, stream(), s.dmesh(), s.apc.p, s.ctrl.rho, s.af.p);
  D.correct(d_pp);
}

// ---------------------------------------------------------- wall zones
static void ensure_walls(System &s) {
  if (s.walls_ready) return;
  s.wall_faces.clear();
  for (int f = 0; f < s.nf; f++) {
    const int g = s.fgrp[f];
    if (g > 0 && s.groups[g - 1].btype == MS_BC_WALL) s.wall_faces.push_back(f);
  }
  if (!s.wall_faces.empty()) {   // face order, device face ids
    std::vector<int> dev(s.wall_faces.size());
    for (size_t k = 0; k < dev.size(); k++) dev[k] = s.fo2d[s.wall_faces[k]];
    s.d_wall_faces.upload(dev);
  }
  s.walls_ready = true;
}

void free_stream(const System &s, double *u_inf, double *p_inf) {
  *u_inf = 0.0;
  *p_inf = 0.0;
  for (int g = 0; g < s.nbc; g++)
    if (s.groups[g].btype == MS_BC_FARFIELD) {
      *u_inf = norm2_ref(s.groups[g].uvel, 3);
      *p_inf = s.groups[g].pval;
      return;
    }
}

void wall_forces(System &s, double *out) {
  ensure_walls(s);
  for (int k = 0; k < 14; k++) out[k] = 0.0;
  double u_inf, p_inf;
  free_stream(s, &u_inf, &p_inf);
  out[5] = u_inf;
  out[7] = p_inf;
  const int nw = (int)s.wall_faces.size();
  if (nw == 0) return;
  double xmin = DBL_MAX, xmax = -DBL_MAX;
  for (int f : s.wall_faces) {
    xmin = std::min(xmin, s.xf[3 * (size_t)f]);
    xmax = std::max(xmax, s.xf[3 * (size_t)f]);
  }
  const double a_ref = (xmax - xmin) * s.zspan;
  hipLaunchKernelGGL(k_wall_forces, gc(nw), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), s.d_wall_faces.p, nw,
                     s.ctrl.mu, p_inf, s.part.p);
  hipLaunchKernelGGL((k_finish<OpSum, 6, double, FinStore>), dim3(1), dim3(kFinThreads), 0, stream(), nblocks(nw),
                     s.part.p, FinStore{s.tmpc.p, 6});
  MS_CHECK_LAUNCH();
  double f[6];
  MS_HIP(hipMemcpyAsync(f, s.tmpc.p, 6 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  MS_HIP(hipStreamSynchronize(stream()));
  out[0] = f[0] + f[3];
  out[1] = f[1] + f[4];
  out[2] = f[2] + f[5];
  out[6] = a_ref;
  for (int k = 0; k < 6; k++) out[8 + k] = f[k];
  if (u_inf > 0.0 && a_ref > 0.0) {
    const double q = 0.5 * s.ctrl.rho * (u_inf * u_inf);
    out[3] = out[0] / (q * a_ref);
    out[4] = out[1] / (q * a_ref);
  }
}

void wall_samples(System &s, std::vector<double> &pu) {
  ensure_walls(s);
  const int nw = (int)s.wall_faces.size();
  pu.assign(4 * (size_t)nw, 0.0);
  if (nw == 0) return;
  DBuf<double> d(4 * (size_t)nw);
  hipLaunchKernelGGL(k_wall_samples, gc(nw), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), s.d_wall_faces.p,
                     nw, d.p);
  MS_CHECK_LAUNCH();
  d.download(pu.data(), pu.size());
  MS_HIP(hipStreamSynchronize(stream()));
}

double wake_end(System &s, double xcyl) {
  hipLaunchKernelGGL(k_wake_end, gc(s.nc), dim3(kBlock), 0, stream(), s.dmesh(), s.dfields(), xcyl, s.part.p);
  hipLaunchKernelGGL((k_finish<OpMax, 1, double, FinStore>), dim3(1), dim3(kFinThreads), 0, stream(), nblocks(s.nc),
                     s.part.p, FinStore{s.tmpc.p, 1});
  MS_CHECK_LAUNCH();
  double v = 0.0;
  MS_HIP(hipMemcpyAsync(&v, s.tmpc.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
  MS_HIP(hipStreamSynchronize(stream()));
  return v;
}

}  // namespace ms
