// Vector-Jacobian product of the interior residual on gfx950.
//
// Given a cotangent w(nshg,4) (same layout as res), computes
//   ybar  = w^T d res / d y      (nshg,4), y = (u1,u2,u3,p)
//   acbar = w^T d res / d ac     (nshg,4), ac = (a1,a2,a3, unused), optional
//   mubar = w^T d res / d mu     (scalar)
// for res the interior ElmGMR residual that crimson_form_interior returns
// (flux pass, e3 with itau = 0 and iconvflow = 2, scatter; before bc3Res).
// The derivative is the exact one of the formulation as written (no lazy
// tangent): tauM and tauC through u (e3stab.f90:108-133), tauBar and ubar
// through the strong residual L (e3stab.f90:356-374), and the reconstructed
// flux q through qpbc's lumped projection of mu (grad u + grad u^T)
// (e3q.f90:100-145, qpbc.f90:67-71).  The acceleration enters the momentum
// residual's rNa and the strong residual L (e3res.f90:775-789), both with
// factor rho.
//
// Reverse mode written by hand and deterministic (no atomics):
//   vjp_e3_elem    element: recompute each quadrature point forward, then
//                  back-propagate the element's residual cotangent; writes
//                  ybar and qbar for its four nodes and its mubar part
//   vjp_q_node     node: qbar gathered over the node's incidences in element
//                  order, times 1/rmass (the adjoint of q = qres / rmass)
//   vjp_flux_elem  element: the adjoint of the lumped projection, qdi's
//                  cotangent sum_a m_a qbar_a / rmass_a, back to u
//   vjp_node       node: ybar gathered over the incidences in element order
//   mu_sum         the mubar parts summed in a fixed order

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "form_common.hpp"

namespace {

constexpr int kVRec = 56;  // ybar (4 nodes x 4) | qbar (4 nodes x 6) | mubar | acbar (4 nodes x 3)
constexpr int V_Y = 0, V_Q = 16, V_MU = 40, V_A = 41;
constexpr int kVRec2 = 16;  // ubar (4 nodes x 3) | mubar
constexpr int V2_MU = 12;

__global__ void __launch_bounds__(256) vjp_e3_elem(int nel, const int *__restrict__ ien,
                                                   const double *__restrict__ node, const double *__restrict__ w,
                                                   int nshg, FormConst c0, Ens en, double *__restrict__ vrec) {
  const int e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  FormConst c = c0;
  c.mu = ens_mu(en, e, c0.mu);
  const size_t N = nshg;
  int nd[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) nd[a] = ien[a * (size_t)nel + e];
  double nl[4][kNRec];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double2 *r = reinterpret_cast<const double2 *>(node + (size_t)nd[a] * kNRec);
#pragma unroll
    for (int k = 0; k < kNRec / 2; ++k) {
      const double2 v = r[k];
      nl[a][2 * k] = v.x;
      nl[a][2 * k + 1] = v.y;
    }
  }
  double g[4][3], W, gg[6];
  {
    double xl[4][3], dx[3][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) xl[a][i] = nl[a][N_X + i];
    p1_metric(xl, dx, W, g);
    gijd_tet(dx, gg);
  }
  double gu[3][3], gp[3], divq[3] = {0.0, 0.0, 0.0};  // gu[j][i] = du_i/dx_j
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
      gu[j][i] = g[0][j] * nl[0][N_U + i] + g[1][j] * nl[1][N_U + i] + g[2][j] * nl[2][N_U + i] +
                 g[3][j] * nl[3][N_U + i];
    gp[j] = g[0][j] * nl[0][N_P] + g[1][j] * nl[1][N_P] + g[2][j] * nl[2][N_P] + g[3][j] * nl[3][N_P];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      divq[i] += g[a][0] * nl[a][N_Q + qsym(i)] + g[a][1] * nl[a][N_Q + qsym(i + 3)] +
                 g[a][2] * nl[a][N_Q + qsym(i + 6)];
  const double rho = c.rho, mu = c.mu, rhoinv = 1.0 / rho;
  const double rnu = mu * rhoinv;
  const double dts = c.Dtgl * c.dtsfct;
  const double GG = gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2] +
                    2.0 * (gg[3] * gg[3] + gg[4] * gg[4] + gg[5] * gg[5]);
  const double tau0 = (2.0 * dts) * (2.0 * dts);
  const double visc = 36.0 * (rnu * rnu) * GG;
  const double trg = gg[0] + gg[1] + gg[2];
  const double divu = gu[0][0] + gu[1][1] + gu[2][2];
  const double ff = c.taucfct / c.dtsfct;

  // cotangents of the element sums: rl_a,i = -W (RNa_ai + g_a . RGN_i:),
  // rl_a,4 = W g_a . SU
  double wl[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int k = 0; k < 4; ++k) wl[a][k] = w[nd[a] + N * k];
  double RGb[3][3], SUb[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
      RGb[i][j] = -W * (wl[0][i] * g[0][j] + wl[1][i] * g[1][j] + wl[2][i] * g[2][j] + wl[3][i] * g[3][j]);
    SUb[i] = W * (wl[0][3] * g[0][i] + wl[1][3] * g[1][i] + wl[2][3] * g[2][i] + wl[3][3] * g[3][i]);
  }

  double gub[3][3], gpb[3] = {0, 0, 0}, divqb[3] = {0, 0, 0}, divub = 0.0, mub = 0.0;
  double ulb[4][3], alb[4][3], plb[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gub[i][j] = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) ulb[a][i] = alb[a][i] = 0.0;

#pragma unroll
  for (int qp = 0; qp < 4; ++qp) {
    const double Nq[4] = {nq(qp, 0), nq(qp, 1), nq(qp, 2), nq(qp, 3)};
    // forward (as e3_elem)
    double u[3], aq[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      u[i] = Nq[0] * nl[0][N_U + i] + Nq[1] * nl[1][N_U + i] + Nq[2] * nl[2][N_U + i] + Nq[3] * nl[3][N_U + i];
      aq[i] = Nq[0] * nl[0][N_A + i] + Nq[1] * nl[1][N_A + i] + Nq[2] * nl[2][N_A + i] + Nq[3] * nl[3][N_A + i];
    }
    double rL[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      rL[i] = (aq[i] + u[0] * gu[0][i] + u[1] * gu[1][i] + u[2] * gu[2][i]) * rho + gp[i] - divq[i];
    double Gu[3], GL[3];
    Gu[0] = gg[0] * u[0] + gg[3] * u[1] + gg[5] * u[2];
    Gu[1] = gg[3] * u[0] + gg[1] * u[1] + gg[4] * u[2];
    Gu[2] = gg[5] * u[0] + gg[4] * u[1] + gg[2] * u[2];
    GL[0] = gg[0] * rL[0] + gg[3] * rL[1] + gg[5] * rL[2];
    GL[1] = gg[3] * rL[0] + gg[1] * rL[1] + gg[4] * rL[2];
    GL[2] = gg[5] * rL[0] + gg[4] * rL[1] + gg[2] * rL[2];
    const double uGu = u[0] * Gu[0] + u[1] * Gu[1] + u[2] * Gu[2];
    const double fact = sqrt((tau0 + uGu) + visc);
    const double tauC = rho * 0.125 * fact / trg * ff;
    const double tauM = 1.0 / fact;
    const double sL = rL[0] * GL[0] + rL[1] * GL[1] + rL[2] * GL[2];
    const double tauBar = (sL != 0.0) ? tauM / sqrt(sL) : 0.0;
    double ub[3], lgr[3], lg[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ub[i] = u[i] - tauM * rL[i] * rhoinv;
      lgr[i] = rL[0] * gu[0][i] + rL[1] * gu[1][i] + rL[2] * gu[2][i];
      lg[i] = tauBar * lgr[i];
    }
    // backward: seeds of this quadrature point
    double rnab[3], ubb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      rnab[i] = -W * (Nq[0] * wl[0][i] + Nq[1] * wl[1][i] + Nq[2] * wl[2][i] + Nq[3] * wl[3][i]);
      ubb[i] = SUb[i];
    }
    // rna_i = (aq_i + ub . grad u_i) rho
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        ubb[j] += rho * rnab[i] * gu[j][i];
        gub[j][i] += rho * rnab[i] * ub[j];
      }
    // v_ij = mu (gu_ij + gu_ji) + tauM L_i u_j + lg_i L_j + delta_ij (-p + tauC div u)
    double tMb = 0.0, tCb = 0.0, tBb = 0.0, tmpb = 0.0, rLb[3] = {0, 0, 0}, uqb[3] = {0, 0, 0}, lgb[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double vb = RGb[i][j];
        gub[i][j] += mu * vb;
        gub[j][i] += mu * vb;
        mub += vb * (gu[i][j] + gu[j][i]);
        tMb += vb * rL[i] * u[j];
        rLb[i] += vb * tauM * u[j];
        uqb[j] += vb * tauM * rL[i];
        lgb[i] += vb * rL[j];
        rLb[j] += vb * lg[i];
        if (i == j) tmpb += vb;
      }
    const double pqb = -tmpb;
    tCb += tmpb * divu;
    divub += tmpb * tauC;
    // lg_i = tauBar (L . grad u_i)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      tBb += lgb[i] * lgr[i];
      const double lgrb = tauBar * lgb[i];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        rLb[k] += lgrb * gu[k][i];
        gub[k][i] += lgrb * rL[k];
      }
    }
    // ubar = u - tauM L / rho
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      uqb[i] += ubb[i];
      tMb -= ubb[i] * rL[i] * rhoinv;
      rLb[i] -= ubb[i] * tauM * rhoinv;
    }
    // tauBar = tauM (L G L)^(-1/2)
    if (sL != 0.0) {
      const double isq = 1.0 / sqrt(sL);
      tMb += tBb * isq;
      const double sLb = -0.5 * tBb * tauM * isq / sL;
#pragma unroll
      for (int k = 0; k < 3; ++k) rLb[k] += 2.0 * sLb * GL[k];
    }
    // tauM = 1 / fact, tauC = rho fact / (8 tr G) ff, fact = sqrt(tau0 + uGu + 36 nu^2 G:G)
    const double factb = -tMb * tauM * tauM + tCb * (rho * 0.125 / trg * ff);
    const double Sb = 0.5 * factb / fact;
#pragma unroll
    for (int k = 0; k < 3; ++k) uqb[k] += 2.0 * Sb * Gu[k];
    mub += Sb * (36.0 * 2.0 * rnu * rhoinv * GG);
    // L_i = (a_i + u . grad u_i) rho + dp/dx_i - (div q)_i
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        uqb[j] += rho * rLb[i] * gu[j][i];
        gub[j][i] += rho * rLb[i] * u[j];
      }
      gpb[i] += rLb[i];
      divqb[i] -= rLb[i];
    }
    // the acceleration enters rna_i and L_i, each times rho
    double aqb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) aqb[i] = rho * (rnab[i] + rLb[i]);
    // u, p and a at the quadrature point
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        ulb[a][i] += Nq[a] * uqb[i];
        alb[a][i] += Nq[a] * aqb[i];
      }
      plb[a] += Nq[a] * pqb;
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) gub[i][i] += divub;
  double *o = vrec + (size_t)e * kVRec;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
      o[V_Y + 4 * a + i] = ulb[a][i] + g[a][0] * gub[0][i] + g[a][1] * gub[1][i] + g[a][2] * gub[2][i];
    o[V_Y + 4 * a + 3] = plb[a] + g[a][0] * gpb[0] + g[a][1] * gpb[1] + g[a][2] * gpb[2];
    double qb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) qb[qsym(i + 3 * j)] += g[a][j] * divqb[i];
#pragma unroll
    for (int s = 0; s < 6; ++s) o[V_Q + 6 * a + s] = qb[s];
#pragma unroll
    for (int i = 0; i < 3; ++i) o[V_A + 3 * a + i] = alb[a][i];
  }
  o[V_MU] = mub;
}

// qbar at the nodes: sum over the node's incidences (element order, 4 lanes
// and a fixed shuffle tree), times 1/rmass.
__global__ void __launch_bounds__(256) vjp_q_node(int nshg, const int *__restrict__ inc_ptr,
                                                  const int *__restrict__ inc, const double *__restrict__ vrec,
                                                  const double *__restrict__ invm, double *__restrict__ qbar) {
  const int gid = xcd_block() * blockDim.x + threadIdx.x;
  const int i = gid / kNodeLanes, lane = gid % kNodeLanes;
  double s[6] = {0, 0, 0, 0, 0, 0};
  if (i < nshg)
    for (int p = inc_ptr[i] + lane; p < inc_ptr[i + 1]; p += kNodeLanes) {
      const int ea = inc[p];
      const double *r = vrec + (size_t)(ea >> 2) * kVRec + V_Q + 6 * (ea & 3);
#pragma unroll
      for (int k = 0; k < 6; ++k) s[k] += r[k];
    }
#pragma unroll
  for (int off = 1; off < kNodeLanes; off <<= 1)
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] += __shfl_xor(s[k], off, kNodeLanes);
  if (i >= nshg || lane != 0) return;
  const double m = invm[i];
#pragma unroll
  for (int k = 0; k < 6; ++k) qbar[(size_t)i * 6 + k] = m * s[k];
}

// Adjoint of the lumped projection: the element's viscous flux qdi enters
// q at its nodes with weight m_a / rmass; qdi = mu (grad u + grad u^T)
// (elem_flux's slots 11,12,13,22,23,33).
__global__ void __launch_bounds__(256) vjp_flux_elem(int nel, const int *__restrict__ ien,
                                                     const double *__restrict__ node,
                                                     const double *__restrict__ qbar, double mu0, Ens en,
                                                     double *__restrict__ vrec2) {
  const int e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  const double mu = ens_mu(en, e, mu0);
  int nd[4];
  double xl[4][3], ul[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    nd[a] = ien[a * (size_t)nel + e];
    const double *r = node + (size_t)nd[a] * kNRec;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xl[a][i] = r[N_X + i];
      ul[a][i] = r[N_U + i];
    }
  }
  double dx[3][3], W, g[4][3];
  p1_metric(xl, dx, W, g);
  double gu[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      gu[j][i] = g[0][j] * ul[0][i] + g[1][j] * ul[1][i] + g[2][j] * ul[2][i] + g[3][j] * ul[3][i];
  double qd[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double *qb = qbar + (size_t)nd[a] * 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) qd[k] += flux_weight(W, a) * qb[k];
  }
  double gub[3][3];
  gub[0][0] = 2.0 * mu * qd[0];
  gub[0][1] = mu * qd[1];
  gub[1][0] = mu * qd[1];
  gub[0][2] = mu * qd[2];
  gub[2][0] = mu * qd[2];
  gub[1][1] = 2.0 * mu * qd[3];
  gub[1][2] = mu * qd[4];
  gub[2][1] = mu * qd[4];
  gub[2][2] = 2.0 * mu * qd[5];
  const double mub = 2.0 * gu[0][0] * qd[0] + (gu[0][1] + gu[1][0]) * qd[1] + (gu[0][2] + gu[2][0]) * qd[2] +
                     2.0 * gu[1][1] * qd[3] + (gu[1][2] + gu[2][1]) * qd[4] + 2.0 * gu[2][2] * qd[5];
  double *o = vrec2 + (size_t)e * kVRec2;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) o[3 * a + i] = g[a][0] * gub[0][i] + g[a][1] * gub[1][i] + g[a][2] * gub[2][i];
  o[V2_MU] = mub;
}

// ybar (and, when acbar is given, acbar) at the nodes, (nshg,4)
// column-major like res; acbar's fourth column is zero.
__global__ void __launch_bounds__(256) vjp_node(int nshg, const int *__restrict__ inc_ptr,
                                                const int *__restrict__ inc, const double *__restrict__ vrec,
                                                const double *__restrict__ vrec2, double *__restrict__ ybar,
                                                double *__restrict__ acbar) {
  const int gid = xcd_block() * blockDim.x + threadIdx.x;
  const int i = gid / kNodeLanes, lane = gid % kNodeLanes;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  if (i < nshg)
    for (int p = inc_ptr[i] + lane; p < inc_ptr[i + 1]; p += kNodeLanes) {
      const int ea = inc[p];
      const int e = ea >> 2, a = ea & 3;
      const double *r = vrec + (size_t)e * kVRec + V_Y + 4 * a;
      const double *r2 = vrec2 + (size_t)e * kVRec2 + 3 * a;
      s[0] += r[0] + r2[0];
      s[1] += r[1] + r2[1];
      s[2] += r[2] + r2[2];
      s[3] += r[3];
      if (acbar) {
        const double *ra = vrec + (size_t)e * kVRec + V_A + 3 * a;
        s[4] += ra[0];
        s[5] += ra[1];
        s[6] += ra[2];
      }
    }
#pragma unroll
  for (int off = 1; off < kNodeLanes; off <<= 1)
#pragma unroll
    for (int k = 0; k < 7; ++k) s[k] += __shfl_xor(s[k], off, kNodeLanes);
  if (i >= nshg || lane != 0) return;
  const size_t N = nshg;
#pragma unroll
  for (int k = 0; k < 4; ++k) ybar[i + N * k] = s[k];
  if (acbar) {
#pragma unroll
    for (int k = 0; k < 3; ++k) acbar[i + N * k] = s[4 + k];
    acbar[i + N * 3] = 0.0;
  }
}

// mubar: per-workgroup sums over fixed element ranges, then one workgroup.
constexpr int kMuChunk = 4096;
__global__ void __launch_bounds__(256) mu_part(int nel, const double *__restrict__ vrec,
                                               const double *__restrict__ vrec2, double *__restrict__ part) {
  __shared__ double sh[256];
  const int e0 = blockIdx.x * kMuChunk;
  double s = 0.0;
  for (int e = e0 + (int)threadIdx.x; e < nel && e < e0 + kMuChunk; e += 256)
    s += vrec[(size_t)e * kVRec + V_MU] + vrec2[(size_t)e * kVRec2 + V2_MU];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(256) mu_final(int n, double *__restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) s += part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[n] = sh[0];
}

}  // namespace

namespace crimson {
void vjp_work_sizes(const crimson_form *f, size_t n[6]) {
  const size_t ne = (size_t)std::max(f->nel, 1), nn = (size_t)f->nshg;
  n[0] = ne * kVRec;
  n[1] = ne * kVRec2;
  n[2] = nn * 6;
  n[3] = nn;
  n[4] = (ne + kMuChunk - 1) / kMuChunk + 1;
  n[5] = form_boundary_vjp_doubles(f);
}

int vjp_run(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x, const double *d_w,
            double *d_ybar, double *d_acbar, double *mubar, const VjpWork &w) {
  if (f->wall) {
    crimson::set_error("crimson_form_vjp_dev: the deformable wall's rows have no adjoint here");
    return CRIMSON_EUNSUPPORTED;
  }
  const size_t nn = (size_t)f->nshg;
  const int nmu = (int)(((size_t)std::max(f->nel, 1) + kMuChunk - 1) / kMuChunk);
  hipStream_t s = f->stream;
  const unsigned T = 256;
  const FormConst c = make_const(f->prm);
  CR_TRY(crimson::form_flux_pass(f, d_y, d_ac, d_x, nullptr, w.invm));
  const unsigned ge = crimson::grid_for(f->nel, T), gn = crimson::grid_for(nn * kNodeLanes, T);
  if (f->nel > 0)
    vjp_e3_elem<<<ge, T, 0, s>>>(f->nel, f->d_ien.p, f->d_node.p, d_w, f->nshg, c, f->ens(), w.vrec);
  vjp_q_node<<<gn, T, 0, s>>>(f->nshg, f->d_inc_ptr.p, f->d_inc.p, w.vrec, w.invm, w.qbar);
  if (f->nel > 0)
    vjp_flux_elem<<<ge, T, 0, s>>>(f->nel, f->d_ien.p, f->d_node.p, w.qbar, c.mu, f->ens(), w.vrec2);
  vjp_node<<<gn, T, 0, s>>>(f->nshg, f->d_inc_ptr.p, f->d_inc.p, w.vrec, w.vrec2, d_ybar, d_acbar);
  double mb = 0.0;
  if (f->nel > 0) {
    mu_part<<<nmu, T, 0, s>>>(f->nel, w.vrec, w.vrec2, w.mupart);
    mu_final<<<1, T, 0, s>>>(nmu, w.mupart);
    CR_HIP(hipGetLastError());
    CR_HIP(hipMemcpyAsync(&mb, w.mupart + nmu, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(s));
  double mbb = 0.0;
  CR_TRY(crimson::form_boundary_vjp(f, d_w, d_ybar, &mbb, w.bvrec));  // the boundary rows (crimson_form_set_boundary)
  if (mubar) *mubar = mb + mbb;
  return CRIMSON_OK;
}

}  // namespace crimson

namespace {
int vjp_entry(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x, const double *d_w,
              double *d_ybar, double *d_acbar, double *mubar) {
  if (!f || !d_y || !d_ac || !d_x || !d_w || !d_ybar) {
    crimson::set_error("crimson_form_vjp_dev: null argument");
    return CRIMSON_EARG;
  }
  size_t n[6];
  crimson::vjp_work_sizes(f, n);
  CR_TRY(f->d_vrec.alloc(n[0]));
  CR_TRY(f->d_vrec2.alloc(n[1]));
  CR_TRY(f->d_qbar.alloc(n[2]));
  CR_TRY(f->d_invm.alloc(n[3]));
  CR_TRY(f->d_mupart.alloc(n[4]));
  const crimson::VjpWork w{f->d_vrec.p, f->d_vrec2.p, f->d_qbar.p, f->d_invm.p, f->d_mupart.p, nullptr};
  return crimson::vjp_run(f, d_y, d_ac, d_x, d_w, d_ybar, d_acbar, mubar, w);
}
}  // namespace

extern "C" int crimson_form_vjp_dev(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x,
                                    const double *d_w, double *d_ybar, double *mubar) {
  return vjp_entry(f, d_y, d_ac, d_x, d_w, d_ybar, nullptr, mubar);
}

extern "C" int crimson_form_vjp_ac_dev(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x,
                                       const double *d_w, double *d_ybar, double *d_acbar, double *mubar) {
  if (!d_acbar) {
    crimson::set_error("crimson_form_vjp_ac_dev: null acbar");
    return CRIMSON_EARG;
  }
  return vjp_entry(f, d_y, d_ac, d_x, d_w, d_ybar, d_acbar, mubar);
}
