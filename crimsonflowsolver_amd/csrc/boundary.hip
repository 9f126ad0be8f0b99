// Boundary elements of ElmGMR's boundary loop (elmgmr.f90:325-460) for P1
// tets: AsBMFG (asbmfg.f90:19-80) gathers each boundary element, e3b
// (e3b.f90:103-560) integrates over its face with e3bvar's quantities
// (e3bvar.f90), and local(..., 'scatter ') adds rl into res.
//
// Supported: rigid walls (ideformwall 0), so the boundary xKebe stays zero
// and the boundary loop adds nothing to lhsK/lhsP; iBCB code bits 0 (mass
// flux), 1 (pressure), 2 (traction) and 4 (vessel wall: no traction term);
// no heart model, no Lagrange multipliers; iconvflow 2; Newtonian mu.
// Boundary rule: symtri with 1 or 3 points (symtri.c:56-60), weights doubled
// (genint.f90:86); on the face N = (r, s, t) of nodes 1-3 and N4 = 0.
//
//   bnd_elem   one thread per boundary element: the four node records of
//              the current formation (x, u, p), the exact P1 gradient of u,
//              the face normal and Jacobian, and rl of the three face nodes
//              (12 doubles) into a boundary record
//   bnd_node   one thread per boundary node: res(node) += rl of its
//              incidences in boundary-element order, after the interior rows
//              are written (the reference adds the boundary loop after the
//              interior loop)
//   bnd_vjp_*  the adjoint of the boundary rows for crimson_form_vjp_dev

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"

#include "form_common.hpp"

namespace {

constexpr double kTriA = 0.3333333333333330000000000000000000;  // symtri.c:20
constexpr double kTriB = 0.666666666666667;                     // symtri.c:21
constexpr double kTriC = 0.166666666666667;                     // symtri.c:22
constexpr int kBV = 17;  // boundary VJP record: (u1,u2,u3,p) cotangents of the 4 nodes, mu cotangent

__device__ inline double tri_pt(int nqb, int q, int a) {
  if (nqb == 1) return kTriA;
  return q == a ? kTriB : kTriC;
}

__global__ void __launch_bounds__(256) bnd_elem(int nelb, const int *__restrict__ ienb,
                                                const int *__restrict__ bcode, const double *__restrict__ bcb,
                                                const int *__restrict__ permit, int nsrf, int nqb, double influx,
                                                double rho, double mu0, Ens en, const double *__restrict__ node,
                                                double *__restrict__ brec, double *__restrict__ bflx) {
  const int e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= nelb) return;
  const double mu = ens_mu(en, e, mu0);
  double xl[4][3], yl[4][4];  // yl(a, (p, u1, u2, u3)) as localy orders it
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double *r = node + (size_t)ienb[a * (size_t)nelb + e] * kNRec;
#pragma unroll
    for (int i = 0; i < 3; ++i) xl[a][i] = r[N_X + i];
    yl[a][0] = r[N_P];
#pragma unroll
    for (int i = 0; i < 3; ++i) yl[a][i + 1] = r[N_U + i];
  }
  // e3bvar.f90: dxdxib from shglb = dN/dxi / 2, its inverse by cofactors,
  // the global gradients gy[j][c] = d(yl c)/dx_j
  constexpr double sg[3][4] = {{0.5, 0.0, 0.0, -0.5}, {0.0, 0.5, 0.0, -0.5}, {0.0, 0.0, 0.5, -0.5}};
  double J[3][3], D[3][3], gl[3][4], gy[3][4];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s = s + xl[a][i] * sg[j][a];
      J[i][j] = s;
    }
  D[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  D[0][1] = J[2][1] * J[0][2] - J[0][1] * J[2][2];
  D[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  const double tdet = 1.0 / (D[0][0] * J[0][0] + D[0][1] * J[1][0] + D[0][2] * J[2][0]);
  D[0][0] *= tdet;
  D[0][1] *= tdet;
  D[0][2] *= tdet;
  D[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * tdet;
  D[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * tdet;
  D[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) * tdet;
  D[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * tdet;
  D[2][1] = (J[2][0] * J[0][1] - J[0][0] * J[2][1]) * tdet;
  D[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * tdet;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 1; c < 4; ++c) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s = s + sg[k][a] * yl[a][c];
      gl[k][c] = s;
    }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int c = 1; c < 4; ++c) gy[j][c] = D[0][j] * gl[0][c] + D[1][j] * gl[1][c] + D[2][j] * gl[2][c];
  // normal from nodes 1-3 and the face Jacobian (e3bvar.f90, lcsyst 1)
  double v1[3], v2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v1[i] = xl[1][i] - xl[0][i];
    v2[i] = xl[2][i] - xl[0][i];
  }
  const double t1 = v1[1] * v2[2] - v2[1] * v1[2];
  const double t2 = v2[0] * v1[2] - v1[0] * v2[2];
  const double t3 = v1[0] * v2[1] - v2[0] * v1[1];
  const double temp = 1.0 / sqrt(t1 * t1 + t2 * t2 + t3 * t3);
  const double bn[3] = {t1 * temp, t2 * temp, t3 * temp};
  const int code = bcode[e], srf = bcode[(size_t)nelb + e];
  const bool flow_ok = !permit || srf < 0 || srf > nsrf || permit[srf] != 0;
  // the viscous traction of the element's constant gradient (e3bvar.f90)
  double tau0[3];
  tau0[0] = bn[0] * 2.0 * mu * gy[0][1] + bn[1] * (mu * (gy[1][1] + gy[0][2])) + bn[2] * (mu * (gy[2][1] + gy[0][3]));
  tau0[1] = bn[0] * (mu * (gy[1][1] + gy[0][2])) + bn[1] * 2.0 * mu * gy[1][2] + bn[2] * (mu * (gy[2][2] + gy[1][3]));
  tau0[2] = bn[0] * (mu * (gy[2][1] + gy[0][3])) + bn[1] * (mu * (gy[2][2] + gy[1][3])) + bn[2] * 2.0 * mu * gy[2][3];
  const double tn = bn[0] * tau0[0] + bn[1] * tau0[1] + bn[2] * tau0[2];
  double rl[3][4];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int k = 0; k < 4; ++k) rl[a][k] = 0.0;
  const double qw = nqb == 1 ? 2.0 * 1.0 : 2.0 * kTriA;
  double fl[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // surface integrals (e3b.f90:216-247)
  for (int q = 0; q < nqb; ++q) {
    const double N[3] = {tri_pt(nqb, q, 0), tri_pt(nqb, q, 1), tri_pt(nqb, q, 2)};
    const double WdetJb = qw / (4.0 * temp);
    double pres = 0.0, u[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pres = pres + N[a] * yl[a][0];
#pragma unroll
      for (int i = 0; i < 3; ++i) u[i] = u[i] + N[a] * yl[a][i + 1];
    }
    double unm = bn[0] * u[0] + bn[1] * u[1] + bn[2] * u[2];
    pres = pres - tn;
    double tau[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) tau[i] = tau0[i] - bn[i] * tn;
    fl[0] = fl[0] + WdetJb;
    fl[1] = fl[1] - WdetJb * unm;
#pragma unroll
    for (int i = 0; i < 3; ++i) fl[2 + i] = fl[2 + i] - (tau[i] - bn[i] * pres) * WdetJb;
    // e3b.f90: influx stabilisation, then the prescribed fluxes
    double unm_copy = 0.0;
    if ((code & 2) && unm < 0.0) unm_copy = unm;
    if (code & 1) {
      unm = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) unm = unm + N[a] * bcb[e + (size_t)nelb * (a + 3 * 0)];
    }
    if ((code & 2) && flow_ok) {
      pres = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) pres = pres + N[a] * bcb[e + (size_t)nelb * (a + 3 * 1)];
    }
    if (code & 4) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        tau[i] = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) tau[i] = tau[i] + N[a] * bcb[e + (size_t)nelb * (a + 3 * (2 + i))];
      }
    }
    if (code & 16) {
      pres = 0.0;
      tau[0] = tau[1] = tau[2] = 0.0;
    }
    double rNa[4];
#pragma unroll
    for (int i = 0; i < 3; ++i) rNa[i] = -tau[i] + bn[i] * pres;
    rNa[3] = unm;
    const double rou_copy = rho * unm_copy;
#pragma unroll
    for (int i = 0; i < 3; ++i) rNa[i] = rNa[i] - influx * rou_copy * u[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) rNa[k] = rNa[k] * WdetJb;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int k = 0; k < 4; ++k) rl[a][k] = rl[a][k] - N[a] * rNa[k];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int k = 0; k < 4; ++k) brec[(size_t)e * 12 + 4 * a + k] = rl[a][k];
#pragma unroll
  for (int k = 0; k < 5; ++k) bflx[(size_t)e * 5 + k] = fl[k];
}

template <int MODE>
__global__ void __launch_bounds__(256) bnd_node(int nbnode, int nshg, const int *__restrict__ bnode,
                                                const int *__restrict__ binc_ptr, const int *__restrict__ binc,
                                                const double *__restrict__ brec, double *__restrict__ res) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbnode) return;
  const int i = bnode[b];
  double r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = MODE == 0 ? res[i + (size_t)nshg * k] : res[4 * (size_t)i + k];
  for (int p = binc_ptr[b]; p < binc_ptr[b + 1]; ++p) {
    const int ea = binc[p];
    const double *rl = brec + (size_t)(ea >> 2) * 12 + 4 * (ea & 3);
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = r[k] + rl[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (MODE == 0)
      res[i + (size_t)nshg * k] = r[k];
    else
      res[4 * (size_t)i + k] = r[k];
  }
}

// Adjoint of bnd_elem: w^T d rl/d (u, p) of the four tet nodes and
// w^T d rl/d mu, for the cotangent w(nshg,4) laid out like res.  Reverse
// mode by hand through the same quantities: the face-point values (p, u,
// unm), the traction of the constant gradient, the code-bit overrides
// (constants carry no derivative) and the influx term's branch.
__global__ void __launch_bounds__(256) bnd_vjp_elem(int nelb, int nshg, const int *__restrict__ ienb,
                                                    const int *__restrict__ bcode, const int *__restrict__ permit,
                                                    int nsrf, int nqb, double influx, double rho, double mu0, Ens en,
                                                    const double *__restrict__ node, const double *__restrict__ w,
                                                    double *__restrict__ vrec) {
  const int e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= nelb) return;
  const double mu = ens_mu(en, e, mu0);
  double xl[4][3], yl[4][4], wl[3][4];
  int g[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    g[a] = ienb[a * (size_t)nelb + e];
    const double *r = node + (size_t)g[a] * kNRec;
#pragma unroll
    for (int i = 0; i < 3; ++i) xl[a][i] = r[N_X + i];
    yl[a][0] = r[N_P];
#pragma unroll
    for (int i = 0; i < 3; ++i) yl[a][i + 1] = r[N_U + i];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int k = 0; k < 4; ++k) wl[a][k] = w[g[a] + (size_t)nshg * k];
  constexpr double sg[3][4] = {{0.5, 0.0, 0.0, -0.5}, {0.0, 0.5, 0.0, -0.5}, {0.0, 0.0, 0.5, -0.5}};
  double J[3][3], D[3][3], gl[3][4], gy[3][4], Gn[4][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s = s + xl[a][i] * sg[j][a];
      J[i][j] = s;
    }
  D[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  D[0][1] = J[2][1] * J[0][2] - J[0][1] * J[2][2];
  D[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  const double tdet = 1.0 / (D[0][0] * J[0][0] + D[0][1] * J[1][0] + D[0][2] * J[2][0]);
  D[0][0] *= tdet;
  D[0][1] *= tdet;
  D[0][2] *= tdet;
  D[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * tdet;
  D[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * tdet;
  D[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) * tdet;
  D[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * tdet;
  D[2][1] = (J[2][0] * J[0][1] - J[0][0] * J[2][1]) * tdet;
  D[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * tdet;
  // Gn[a][j] = d(gy[j][c]) / d(yl[a][c]): the P1 gradient of node a
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) Gn[a][j] = D[0][j] * sg[0][a] + D[1][j] * sg[1][a] + D[2][j] * sg[2][a];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 1; c < 4; ++c) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s = s + sg[k][a] * yl[a][c];
      gl[k][c] = s;
    }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int c = 1; c < 4; ++c) gy[j][c] = D[0][j] * gl[0][c] + D[1][j] * gl[1][c] + D[2][j] * gl[2][c];
  double v1[3], v2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v1[i] = xl[1][i] - xl[0][i];
    v2[i] = xl[2][i] - xl[0][i];
  }
  const double t1 = v1[1] * v2[2] - v2[1] * v1[2];
  const double t2 = v2[0] * v1[2] - v1[0] * v2[2];
  const double t3 = v1[0] * v2[1] - v2[0] * v1[1];
  const double temp = 1.0 / sqrt(t1 * t1 + t2 * t2 + t3 * t3);
  const double bn[3] = {t1 * temp, t2 * temp, t3 * temp};
  const int code = bcode[e], srf = bcode[(size_t)nelb + e];
  const bool flow_ok = !permit || srf < 0 || srf > nsrf || permit[srf] != 0;
  // (grad u + grad u^T) n, so that tau0 = mu sym_n
  double symn[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) s = s + (gy[j][k + 1] + gy[k][j + 1]) * bn[j];
    symn[k] = s;
  }
  const bool p_free = !(code & 16) && !((code & 2) && flow_ok);  // pres = p(q) - tn
  const bool t_free = !(code & 16) && !(code & 4);               // tau = tau0 - n tn
  double ub[4][3], pb[3], tnb = 0.0, t0b[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 4; ++a) ub[a][0] = ub[a][1] = ub[a][2] = 0.0;
  pb[0] = pb[1] = pb[2] = 0.0;
  const double qw = nqb == 1 ? 2.0 * 1.0 : 2.0 * kTriA;
  for (int q = 0; q < nqb; ++q) {
    const double N[3] = {tri_pt(nqb, q, 0), tri_pt(nqb, q, 1), tri_pt(nqb, q, 2)};
    const double W = qw / (4.0 * temp);
    double u[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) u[i] = u[i] + N[a] * yl[a][i + 1];
    const double unm = bn[0] * u[0] + bn[1] * u[1] + bn[2] * u[2];
    double gk[4];  // d J / d raw_k at this point
#pragma unroll
    for (int k = 0; k < 4; ++k) gk[k] = -W * (N[0] * wl[0][k] + N[1] * wl[1][k] + N[2] * wl[2][k]);
    double uq[3] = {0.0, 0.0, 0.0}, pq = 0.0;
    if (!(code & 1))
#pragma unroll
      for (int j = 0; j < 3; ++j) uq[j] += gk[3] * bn[j];
    const double gn = gk[0] * bn[0] + gk[1] * bn[1] + gk[2] * bn[2];
    if (p_free) {
      pq += gn;
      tnb -= gn;
    }
    if (t_free) {
#pragma unroll
      for (int i = 0; i < 3; ++i) t0b[i] -= gk[i];
      tnb += gn;
    }
    if ((code & 2) && unm < 0.0) {
      const double cr = influx * rho;
      const double gu = gk[0] * u[0] + gk[1] * u[1] + gk[2] * u[2];
#pragma unroll
      for (int j = 0; j < 3; ++j) uq[j] -= cr * (bn[j] * gu + unm * gk[j]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pb[a] += N[a] * pq;
#pragma unroll
      for (int j = 0; j < 3; ++j) ub[a][j] += N[a] * uq[j];
    }
  }
  // tn = n . tau0, tau0 = mu (grad u + grad u^T) n
#pragma unroll
  for (int i = 0; i < 3; ++i) t0b[i] += tnb * bn[i];
  const double mub = t0b[0] * symn[0] + t0b[1] * symn[1] + t0b[2] * symn[2];
  // d tau0_k / d (du_k'/dx_j): mu (delta_kk' n_j + delta_jk' n_k)
#pragma unroll
  for (int kp = 0; kp < 3; ++kp)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double db = mu * (t0b[kp] * bn[j] + t0b[j] * bn[kp]);
#pragma unroll
      for (int a = 0; a < 4; ++a) ub[a][kp] += db * Gn[a][j];
    }
  double *o = vrec + (size_t)e * kBV;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int i = 0; i < 3; ++i) o[4 * a + i] = ub[a][i];
    o[4 * a + 3] = a < 3 ? pb[a] : 0.0;
  }
  o[16] = mub;
}

// ybar(node, :) += the boundary records of its incidences (all four tet
// nodes), in boundary-element order
__global__ void __launch_bounds__(256) bnd_vjp_node(int nvn, int nshg, const int *__restrict__ vnode,
                                                    const int *__restrict__ vptr, const int *__restrict__ vinc,
                                                    const double *__restrict__ vrec, double *__restrict__ ybar) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nvn) return;
  const int i = vnode[b];
  double r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = ybar[i + (size_t)nshg * k];
  for (int p = vptr[b]; p < vptr[b + 1]; ++p) {
    const int ea = vinc[p];
    const double *v = vrec + (size_t)(ea >> 2) * kBV + 4 * (ea & 3);
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = r[k] + v[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) ybar[i + (size_t)nshg * k] = r[k];
}

// sum of the boundary elements' mu cotangents, one workgroup, fixed order
__global__ void __launch_bounds__(256) bnd_vjp_mu(int nelb, const double *__restrict__ vrec,
                                                  double *__restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int e = threadIdx.x; e < nelb; e += 256) s += vrec[(size_t)e * kBV + 16];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

}  // namespace

// flxID(k, s) = sum over the boundary elements of srfID +-s (sorted per
// surface, in element order: selem[sptr[s] .. sptr[s+1])) of their
// integrals; one workgroup per surface, fixed order.
__global__ void __launch_bounds__(256) bnd_flux_sum(int nflx, const int *__restrict__ sptr,
                                                    const int *__restrict__ selem, const double *__restrict__ bflx,
                                                    double *__restrict__ out) {
  __shared__ double sh[5][4];
  const int s = blockIdx.x;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int p = sptr[s] + threadIdx.x; p < sptr[s + 1]; p += 256) {
    const double *r = bflx + (size_t)selem[p] * 5;
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[k] += r[k];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x / 64] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5) out[5 * (size_t)s + threadIdx.x] = (sh[threadIdx.x][0] + sh[threadIdx.x][1]) +
                                                           (sh[threadIdx.x][2] + sh[threadIdx.x][3]);
}

// The ensemble instance of each boundary element: that of the first element
// holding its first node (instances have disjoint nodes).
__global__ void bnd_inst_kernel(int nelb, const int *__restrict__ ienb, const int *__restrict__ inc_ptr,
                                const int *__restrict__ inc, const unsigned short *__restrict__ inst,
                                unsigned short *__restrict__ binst) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelb) return;
  const int n = ienb[e], p = inc_ptr[n];
  binst[e] = p < inc_ptr[n + 1] ? inst[inc[p] >> 2] : 0;  // a node of no element: instance 0
}

namespace crimson {
int form_boundary_instances(crimson_form *f) {
  if (f->nelb == 0 || f->ninst == 0) return CRIMSON_OK;
  CR_TRY(f->d_binst.alloc(f->nelb));
  bnd_inst_kernel<<<grid_for(f->nelb, 256), 256, 0, f->stream>>>(f->nelb, f->d_ienb.p, f->d_inc_ptr.p, f->d_inc.p,
                                                                 f->d_inst.p, f->d_binst.p);
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(f->stream));
  return CRIMSON_OK;
}
int form_boundary_pass(crimson_form *f, double *d_res, int mode) {
  if (f->nelb == 0) return CRIMSON_OK;
  hipStream_t s = f->stream;
  bnd_elem<<<grid_for(f->nelb, 256), 256, 0, s>>>(f->nelb, f->d_ienb.p, f->d_bcode.p, f->d_bcb.p,
                                                  f->nsrf >= 0 ? f->d_permit.p : nullptr, f->nsrf, f->nqb,
                                                  f->influx, f->prm.rho, f->prm.mu, f->bens(), f->d_node.p,
                                                  f->d_brec.p, f->d_bflx.p);
  CR_TRY(form_wall_rows(f));  // the wall's rows join each face's rl (e3b.f90:488-575)
  if (mode == 0)
    bnd_node<0><<<grid_for(f->nbnode, 256), 256, 0, s>>>(f->nbnode, f->nshg, f->d_bnode.p, f->d_binc_ptr.p,
                                                         f->d_binc.p, f->d_brec.p, d_res);
  else
    bnd_node<1><<<grid_for(f->nbnode, 256), 256, 0, s>>>(f->nbnode, f->nshg, f->d_bnode.p, f->d_binc_ptr.p,
                                                         f->d_binc.p, f->d_brec.p, d_res);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}
size_t form_boundary_vjp_doubles(const crimson_form *f) { return f->nelb ? (size_t)kBV * f->nelb + 1 : 0; }
int form_boundary_vjp(crimson_form *f, const double *d_w, double *d_ybar, double *mubar, double *d_bvrec) {
  *mubar = 0.0;
  if (f->nelb == 0) return CRIMSON_OK;
  hipStream_t s = f->stream;
  if (!d_bvrec) {
    CR_TRY(f->d_bvrec.alloc(form_boundary_vjp_doubles(f)));
    d_bvrec = f->d_bvrec.p;
  }
  bnd_vjp_elem<<<grid_for(f->nelb, 256), 256, 0, s>>>(f->nelb, f->nshg, f->d_ienb.p, f->d_bcode.p,
                                                      f->nsrf >= 0 ? f->d_permit.p : nullptr, f->nsrf, f->nqb,
                                                      f->influx, f->prm.rho, f->prm.mu, f->bens(), f->d_node.p, d_w,
                                                      d_bvrec);
  bnd_vjp_node<<<grid_for(f->nvnode, 256), 256, 0, s>>>(f->nvnode, f->nshg, f->d_vnode.p, f->d_vptr.p,
                                                        f->d_vinc.p, d_bvrec, d_ybar);
  double *d_mu = d_bvrec + (size_t)kBV * f->nelb;
  bnd_vjp_mu<<<1, 256, 0, s>>>(f->nelb, d_bvrec, d_mu);
  CR_HIP(hipGetLastError());
  CR_HIP(hipMemcpyAsync(mubar, d_mu, sizeof(double), hipMemcpyDeviceToHost, s));
  CR_HIP(hipStreamSynchronize(s));
  return CRIMSON_OK;
}
}  // namespace crimson

extern "C" int crimson_form_set_boundary(crimson_form *f, int nelb, const int *ienb, const int *iBCB,
                                         const double *BCB, int ndBCB, int nqb, double influx,
                                         const int *permit, int nsrf) {
  if (!f) return crimson::set_error("crimson_form_set_boundary: null handle"), CRIMSON_EARG;
  if (f->wall) {  // the wall's faces index the old boundary
    f->wall = false;
    f->nwf = f->nwn = f->nwk = 0;
    f->wall_gen++;
  }
  if (nelb == 0) {
    f->nelb = f->nbnode = 0;
    if (f->npartial > 0) {
      f->h_nbnd.assign(f->nshg, 0);
      return crimson::form_bc3lhs_table(f);
    }
    return CRIMSON_OK;
  }
  if (nelb < 0 || !ienb || !iBCB || !BCB || ndBCB < 2 || (nqb != 1 && nqb != 3) || (permit && nsrf < 0))
    return crimson::set_error("crimson_form_set_boundary: bad arguments"), CRIMSON_EARG;
  std::vector<int> ib((size_t)4 * nelb), code((size_t)2 * nelb);
  std::vector<double> bcb((size_t)nelb * 15, 0.0);
  for (int e = 0; e < nelb; ++e) {
    for (int a = 0; a < 4; ++a) {
      const int g = ienb[e + (size_t)nelb * a] - 1;
      if (g < 0 || g >= f->nshg || g >= f->numnp)
        return crimson::set_error("crimson_form_set_boundary: ienb(%d,%d) out of range", e + 1, a + 1), CRIMSON_EARG;
      ib[(size_t)a * nelb + e] = g;
    }
    const int c = iBCB[e];
    if ((c & 4) && ndBCB < 5)
      return crimson::set_error("crimson_form_set_boundary: traction bit set with ndBCB < 5"), CRIMSON_EARG;
    code[e] = c;
    code[(size_t)nelb + e] = iBCB[e + (size_t)nelb];
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < std::min(ndBCB, 5); ++k)
        bcb[e + (size_t)nelb * (a + 3 * k)] = BCB[e + (size_t)nelb * (a + 3 * k)];
  }
  // boundary nodes and their incidences (element, face node) in element order
  std::vector<int> cnt(f->nshg, 0);
  for (int e = 0; e < nelb; ++e)
    for (int a = 0; a < 3; ++a) cnt[ib[(size_t)a * nelb + e]]++;
  std::vector<int> bnode, bptr(1, 0), pos(f->nshg, -1);
  for (int i = 0; i < f->nshg; ++i)
    if (cnt[i]) {
      pos[i] = (int)bnode.size();
      bnode.push_back(i);
      bptr.push_back(bptr.back() + cnt[i]);
    }
  std::vector<int> binc(bptr.back()), fill(bptr.begin(), bptr.end() - 1);
  for (int e = 0; e < nelb; ++e)
    for (int a = 0; a < 3; ++a) {
      const int b = pos[ib[(size_t)a * nelb + e]];
      binc[fill[b]++] = e << 2 | a;
    }
  // the adjoint's incidences: all four nodes of each boundary element
  std::vector<int> vcnt(f->nshg, 0);
  for (int e = 0; e < nelb; ++e)
    for (int a = 0; a < 4; ++a) vcnt[ib[(size_t)a * nelb + e]]++;
  std::vector<int> vnode, vptr(1, 0), vpos(f->nshg, -1);
  for (int i = 0; i < f->nshg; ++i)
    if (vcnt[i]) {
      vpos[i] = (int)vnode.size();
      vnode.push_back(i);
      vptr.push_back(vptr.back() + vcnt[i]);
    }
  std::vector<int> vinc(vptr.back()), vfill(vptr.begin(), vptr.end() - 1);
  for (int e = 0; e < nelb; ++e)
    for (int a = 0; a < 4; ++a) {
      const int b = vpos[ib[(size_t)a * nelb + e]];
      vinc[vfill[b]++] = e << 2 | a;
    }
  if (nelb > (1 << 29)) return crimson::set_error("crimson_form_set_boundary: too many elements"), CRIMSON_EARG;
  // boundary elements by |srfID| (0 holds those without a surface), in element order
  int maxs = 0;
  for (int e = 0; e < nelb; ++e) maxs = std::max(maxs, std::abs(code[(size_t)nelb + e]));
  std::vector<int> sptr(maxs + 2, 0), selem(nelb);
  for (int e = 0; e < nelb; ++e) sptr[std::abs(code[(size_t)nelb + e]) + 1]++;
  for (int k = 0; k <= maxs; ++k) sptr[k + 1] += sptr[k];
  {
    std::vector<int> fillp(sptr.begin(), sptr.end() - 1);
    for (int e = 0; e < nelb; ++e) selem[fillp[std::abs(code[(size_t)nelb + e])]++] = e;
  }
  hipStream_t s = f->stream;
  int rc;
  if ((rc = f->d_ienb.upload(ib.data(), ib.size(), s)) || (rc = f->d_bcode.upload(code.data(), code.size(), s)) ||
      (rc = f->d_bcb.upload(bcb.data(), bcb.size(), s)) || (rc = f->d_brec.alloc((size_t)12 * nelb)) ||
      (rc = f->d_bnode.upload(bnode.data(), bnode.size(), s)) ||
      (rc = f->d_binc_ptr.upload(bptr.data(), bptr.size(), s)) || (rc = f->d_binc.upload(binc.data(), binc.size(), s)) ||
      (rc = f->d_vnode.upload(vnode.data(), vnode.size(), s)) || (rc = f->d_vptr.upload(vptr.data(), vptr.size(), s)) ||
      (rc = f->d_vinc.upload(vinc.data(), vinc.size(), s)) || (rc = f->d_bflx.alloc((size_t)5 * nelb)) ||
      (rc = f->d_sptr.upload(sptr.data(), sptr.size(), s)) || (rc = f->d_selem.upload(selem.data(), selem.size(), s)))
    return rc;
  if (permit) {
    if ((rc = f->d_permit.upload(permit, (size_t)nsrf + 1, s))) return rc;
  }
  CR_HIP(hipStreamSynchronize(s));
  f->nelb = nelb;
  f->nbnode = (int)bnode.size();
  f->nvnode = (int)vnode.size();
  f->nqb = nqb;
  f->influx = influx;
  f->nsrf = permit ? nsrf : -1;
  f->maxsrf = maxs;
  CR_TRY(crimson::form_boundary_instances(f));
  if (f->npartial > 0) {
    // the boundary loop's bc3lhs sets the constrained diagonal entries of
    // each boundary element's partial-BC nodes to one (elmgmr.f90:364-372)
    f->h_nbnd = vcnt;
    return crimson::form_bc3lhs_table(f);
  }
  return CRIMSON_OK;
}

extern "C" int crimson_form_boundary_flux(crimson_form *f, int nflx, double *flxid) {
  if (!f || nflx < 0 || !flxid) return crimson::set_error("crimson_form_boundary_flux: bad arguments"), CRIMSON_EARG;
  for (size_t k = 0; k < 5 * ((size_t)nflx + 1); ++k) flxid[k] = 0.0;
  if (f->nelb == 0) return CRIMSON_OK;
  const int ns = f->maxsrf + 1;
  CR_TRY(f->d_flxout.alloc((size_t)5 * ns));
  bnd_flux_sum<<<ns, 256, 0, f->stream>>>(ns, f->d_sptr.p, f->d_selem.p, f->d_bflx.p, f->d_flxout.p);
  CR_HIP(hipGetLastError());
  std::vector<double> h((size_t)5 * ns);
  CR_HIP(hipMemcpyAsync(h.data(), f->d_flxout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, f->stream));
  CR_HIP(hipStreamSynchronize(f->stream));
  for (int sid = 1; sid <= std::min(nflx, f->maxsrf); ++sid)
    for (int k = 0; k < 5; ++k) flxid[5 * (size_t)sid + k] = h[5 * (size_t)sid + k];
  return CRIMSON_OK;
}

// Staging of crimson_form_boundary_flux_noalloc: bnd_flux_sum's per-surface
// sums, 5 doubles for each |srfID| 0..CRIMSON_FLUX_NOALLOC_MAX_SURFACES-1
// (160 KiB).  Each device holds its own copy for the life of the process.
// bnd_flux_sum writes every column it covers in full, so nothing in it has
// to start at zero; every call on a device shares that device's copy.
static __device__ __align__(256) double scratch[5 * CRIMSON_FLUX_NOALLOC_MAX_SURFACES];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kScratchDevices] = {};

// The device address of `scratch` on `device`, looked up once per device
// with that device current (the host-side name is not a device address).
int flux_scratch(int device, double **out) {
  if (device < 0 || device >= kScratchDevices)
    return crimson::set_error("crimson_form_boundary_flux_noalloc: device %d out of range", device), CRIMSON_EARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    CR_HIP(hipGetDevice(&cur));
    if (cur != device) CR_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) CR_HIP(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return crimson::set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e)), CRIMSON_EHIP;
    g_scratch_ptr[device] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[device];
  return CRIMSON_OK;
}
}  // namespace

extern "C" int crimson_form_boundary_flux_noalloc(crimson_form *f, int nflx, double *flxid) {
  if (!f || nflx < 0 || !flxid)
    return crimson::set_error("crimson_form_boundary_flux_noalloc: bad arguments"), CRIMSON_EARG;
  const int ns = f->maxsrf + 1;
  if (f->nelb > 0 && ns > CRIMSON_FLUX_NOALLOC_MAX_SURFACES)
    return crimson::set_error("crimson_form_boundary_flux_noalloc: srfID up to %d; this entry point holds "
                              "|srfID| < %d (use crimson_form_boundary_flux)",
                              f->maxsrf, CRIMSON_FLUX_NOALLOC_MAX_SURFACES),
           CRIMSON_EARG;
  for (size_t k = 0; k < 5 * ((size_t)nflx + 1); ++k) flxid[k] = 0.0;
  if (f->nelb == 0) return CRIMSON_OK;
  double *d_scratch = nullptr;
  CR_TRY(flux_scratch(f->device, &d_scratch));
  bnd_flux_sum<<<ns, 256, 0, f->stream>>>(ns, f->d_sptr.p, f->d_selem.p, f->d_bflx.p, d_scratch);
  CR_HIP(hipGetLastError());
  const int ncol = std::min(nflx, f->maxsrf);  // columns 1..ncol; column 0 stays zero
  if (ncol > 0)
    CR_HIP(hipMemcpyAsync(flxid + 5, d_scratch + 5, (size_t)5 * ncol * sizeof(double), hipMemcpyDeviceToHost,
                          f->stream));
  CR_HIP(hipStreamSynchronize(f->stream));
  return CRIMSON_OK;
}
