// Stabilized-FEM interior formation for P1 tets on gfx950 (the interior part
// of the reference's ElmGMR, elmgmr.f90:82-255).
//
// Pipeline, all deterministic (no atomics; every sum runs in element order):
//   node_pack   node-parallel: x and u packed as 6 doubles per node
//   flux_node   node-parallel gather over the node's elements in element
//               order, each incidence evaluating its element's metric and
//               viscous flux q_dI (e3q.f90:100-145) from the packed nodes:
//               qres, rmass, q = qres * (1/rmass)  (asiq.f90 + qpbc.f90);
//               writes one 128-byte node record (x, u, p, ac, q)
//   e3_elem     element-parallel e3 (e3.f90 .. e3lhs.f90) on the four node
//               records: all four quadrature points in registers; writes a
//               12-double record per element incidence (e, a) at the
//               incidence's place in the node-major incidence list: node
//               a's row S_a. of the scalar part of the 3x3 velocity blocks,
//               its residual row rl_a, and the element scalars Tc, Mu, Cm,
//               tlW from which every tangent entry is rebuilt
//   asm_row     row-parallel, 16 lanes per CSR row, one lane per block: the
//               row's incidence records read as one contiguous run, the
//               shape gradients rebuilt from the packed nodes, the blocks of
//               the row from the row's element incidences in element order
//               (fillsparseI, fillsparse.f90:29-63), the -G^T column of Val4
//               from the same contributions (solfar.f90:225-235 without its
//               transpose search) and the residual row (asigmr.f90:116)
//   res_node    the residual rows alone when no tangent is formed
//   e3_tangent32, asm_row32   the FP32 tangent of the mixed-precision form
//
// The record factorisation is exact algebra on the reference's
// tangent (e3lhs.f90:85-402, iconvflow = 2 branch :124-129) for P1
// elements, whose metric and shape gradients do not vary over the element:
//   K^{ij}_ab = delta_ij S_ab + Tc g_a,i g_b,j + Mu g_a,j g_b,i
//   S_ab = tsFct sum_q N_a N_b + rho tlW sum_q N_a (ubar_q.g_b)
//          + Mu (g_a.g_b) + rho tlW sum_q tauM_q (u_q.g_a)(u_q.g_b)
//          + tlW sum_q tauBar_q (L_q.g_a)(L_q.g_b)
//   G_ab,i = tlW g_b,i sum_q N_a ,   C_ab = Cm (g_a.g_b)
// with Tc = tlW sum_q tauC_q, Mu = 4 tlW mu, Cm = tlW sum_q tauM_q / rho,
// tlW = alpha_f gamma Delt WdetJ, tsFct = alpha_m WdetJ rho.
// K^{ij} is stored at lhsK(i + 3(j-1)) exactly as e3lhs.f90 places it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"

#include "form_common.hpp"

namespace {

// ---------------------------------------------------------------- flux pass
// x and u of every node packed as kXU doubles: the flux pass gathers a
// node's x and u in three 16-byte loads, the assembly its x in two loads.
__global__ void __launch_bounds__(256) node_pack(int nshg, const double *__restrict__ x, int numnp,
                                                 const double *__restrict__ y, double *__restrict__ xu) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= nshg) return;
  const size_t N = nshg, NP = numnp;
  double2 *o = reinterpret_cast<double2 *>(xu + (size_t)i * kXU);
  o[0] = make_double2(x[i], x[i + NP]);
  o[1] = make_double2(x[i + 2 * NP], y[i]);
  o[2] = make_double2(y[i + N], y[i + 2 * N]);
}

// The coordinates of an element's four nodes (local order) from the packed
// nodes.
__device__ inline void gather_x(const double *__restrict__ xu, int4 n4, double (&xl)[4][3]) {
  const int nd[4] = {n4.x, n4.y, n4.z, n4.w};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double *r = xu + (size_t)nd[a] * kXU;
    const double2 v = *reinterpret_cast<const double2 *>(r);
    xl[a][0] = v.x;
    xl[a][1] = v.y;
    xl[a][2] = r[2];
  }
}

// e3q for one element (e3q.f90:66-145): qdi = mu (grad u + grad u^T) is the
// same at every quadrature point of a P1 element.  Returns WdetJ; qd holds
// the 6 distinct components in slots (11,12,13,22,23,33).
__device__ inline double elem_flux(const double *__restrict__ xu, int4 n4, double mu, double (&qd)[6]) {
  const int nd[4] = {n4.x, n4.y, n4.z, n4.w};
  double xl[4][3], ul[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double2 *r = reinterpret_cast<const double2 *>(xu + (size_t)nd[a] * kXU);
    const double2 v0 = r[0], v1 = r[1], v2 = r[2];
    xl[a][0] = v0.x;
    xl[a][1] = v0.y;
    xl[a][2] = v1.x;
    ul[a][0] = v1.y;
    ul[a][1] = v2.x;
    ul[a][2] = v2.y;
  }
  double dx[3][3], W, g[4][3];
  p1_metric(xl, dx, W, g);
  double gu[3][3];  // gu[j][i] = d u_i / d x_j
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      gu[j][i] = g[0][j] * ul[0][i] + g[1][j] * ul[1][i] + g[2][j] * ul[2][i] + g[3][j] * ul[3][i];
  qd[0] = 2.0 * mu * gu[0][0];
  qd[1] = mu * (gu[0][1] + gu[1][0]);
  qd[2] = mu * (gu[0][2] + gu[2][0]);
  qd[3] = 2.0 * mu * gu[1][1];
  qd[4] = mu * (gu[1][2] + gu[2][1]);
  qd[5] = 2.0 * mu * gu[2][2];
  return W;
}

// qres/rmass gather and qpbc inversion (asiq.f90, qpbc.f90): each of the
// node's incidences (in element order, kNodeLanes lanes) evaluates its
// element's flux from the packed nodes; nothing per element is stored.
// q(nshg,9) column-major in the reference's component order: (1,4,7) =
// (q11,q12,q13), ... so component k = i + 3 (j-1) holds mu(du_i/dx_j +
// du_j/dx_i).  Writes the node record (with the node's x, y, ac) and, when q
// is not null, q(nshg,9).
__global__ void __launch_bounds__(256) flux_node(int nshg, const int *__restrict__ inc_ptr,
                                                 const int *__restrict__ inc, const int4 *__restrict__ ienp,
                                                 const double *__restrict__ xu, double mu0, Ens en,
                                                 const double *__restrict__ x, int numnp,
                                                 const double *__restrict__ y, const double *__restrict__ ac,
                                                 double *__restrict__ node, double *__restrict__ q,
                                                 double *__restrict__ invm) {
  const int gid = xcd_block() * blockDim.x + threadIdx.x;
  const int i = gid / kNodeLanes, lane = gid % kNodeLanes;
  double s[6] = {0, 0, 0, 0, 0, 0}, rm = 0.0;
  if (i < nshg) {
    for (int p = inc_ptr[i] + lane; p < inc_ptr[i + 1]; p += kNodeLanes) {
      const int ea = inc[p];
      double qd[6];
      const double m = flux_weight(elem_flux(xu, ienp[p], ens_mu(en, ea >> 2, mu0), qd), ea & 3);
#pragma unroll
      for (int k = 0; k < 6; ++k) s[k] += m * qd[k];
      rm += m;
    }
  }
  // fixed-order reduction over the node's lanes: (0+1)+(2+3)
#pragma unroll
  for (int off = 1; off < kNodeLanes; off <<= 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] += __shfl_xor(s[k], off, kNodeLanes);
    rm += __shfl_xor(rm, off, kNodeLanes);
  }
  if (i >= nshg) return;
  const double inv = 1.0 / rm;
  const size_t N = nshg, NP = numnp;
  if (invm && lane == 0) invm[i] = inv;
  // the node's record, four doubles per lane
  double v[4];
  if (lane == 0) {
    v[0] = x[i];
    v[1] = x[i + NP];
    v[2] = x[i + 2 * NP];
    v[3] = y[i];
  } else if (lane == 1) {
    v[0] = y[i + N];
    v[1] = y[i + 2 * N];
    v[2] = y[i + 3 * N];
    v[3] = ac[i];
  } else if (lane == 2) {
    v[0] = ac[i + N];
    v[1] = ac[i + 2 * N];
    v[2] = inv * s[0];
    v[3] = inv * s[1];
  } else {
    v[0] = inv * s[2];
    v[1] = inv * s[3];
    v[2] = inv * s[4];
    v[3] = inv * s[5];
  }
  double2 *nr = reinterpret_cast<double2 *>(node + (size_t)i * kNRec + 4 * lane);
  nr[0] = make_double2(v[0], v[1]);
  nr[1] = make_double2(v[2], v[3]);
  if (!q || lane != 0) return;
  q[i + 0 * N] = inv * s[0];
  q[i + 1 * N] = inv * s[1];
  q[i + 2 * N] = inv * s[2];
  q[i + 3 * N] = inv * s[1];
  q[i + 4 * N] = inv * s[3];
  q[i + 5 * N] = inv * s[4];
  q[i + 6 * N] = inv * s[2];
  q[i + 7 * N] = inv * s[4];
  q[i + 8 * N] = inv * s[5];
}

// ---------------------------------------------------------------- e3
// With the tangent, the element records of the workgroup's (consecutive)
// elements are assembled in LDS and stored as one contiguous block with
// coalesced 16-byte stores (written directly, each thread's 256-byte record
// reached HBM as partially filled lines); the scalars (32 bytes per element)
// are stored directly, contiguous across the threads.
constexpr int kE3Threads = 128;
// LDS stride of a staged record: one double more than the record, so that
// the threads' stores at the same offset fall in different banks.
constexpr int kE3Pad = kRec + 1;
template <bool LHS>
__global__ void __launch_bounds__(kE3Threads) __attribute__((amdgpu_waves_per_eu(2))) e3_elem(int nel, const int *__restrict__ ien,
                                                      const double *__restrict__ node, FormConst c0, Ens en,
                                                      double *__restrict__ rec, double *__restrict__ esc) {
  __shared__ double sh_rec[LHS ? kE3Threads * kE3Pad : 1];
  const int e0 = xcd_block() * kE3Threads;
  const int e = e0 + threadIdx.x;
  if (!LHS && e >= nel) return;
  const int ee = e < nel ? e : nel - 1;  // tail threads evaluate a valid element; nothing of theirs is stored
  FormConst c = c0;
  c.mu = ens_mu(en, ee, c0.mu);
  int nd[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) nd[a] = ien[a * (size_t)nel + ee];
  // the four node records (localx/localy gathers, localy.f90:54-64)
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
  // nodal velocity and pressure; gradients (e3ivar.f90:180-198)
  double ul[4][3], pl[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int i = 0; i < 3; ++i) ul[a][i] = nl[a][N_U + i];
    pl[a] = nl[a][N_P];
  }
  double gu[3][3], gp[3];  // gu[j][i] = du_i/dx_j
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
      gu[j][i] = g[0][j] * ul[0][i] + g[1][j] * ul[1][i] + g[2][j] * ul[2][i] + g[3][j] * ul[3][i];
    gp[j] = g[0][j] * pl[0] + g[1][j] * pl[1] + g[2][j] * pl[2] + g[3][j] * pl[3];
  }
  // divergence of the reconstructed flux (e3ivar.f90:355-367)
  double divq[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double qa[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) qa[k] = nl[a][N_Q + qsym(k)];
#pragma unroll
    for (int i = 0; i < 3; ++i) divq[i] += g[a][0] * qa[i] + g[a][1] * qa[i + 3] + g[a][2] * qa[i + 6];
  }
  double al[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) al[a][i] = nl[a][N_A + i];

  const double rho = c.rho, mu = c.mu, rhoinv = 1.0 / rho;
  const double rnu = mu * rhoinv;
  const double dts = c.Dtgl * c.dtsfct;
  const double GG = gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2] +
                    2.0 * (gg[3] * gg[3] + gg[4] * gg[4] + gg[5] * gg[5]);
  const double tau0 = (2.0 * dts) * (2.0 * dts);
  const double visc = 36.0 * (rnu * rnu) * GG;  // fff = 36 for ipord = 1
  const double trg = gg[0] + gg[1] + gg[2];
  const double divu = gu[0][0] + gu[1][1] + gu[2][2];
  const double ff = c.taucfct / c.dtsfct;

  double RNa[4][3], RGN[3][3], SU[3] = {0, 0, 0}, TC = 0.0, TM = 0.0, SUPG[10], UB[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      RNa[a][i] = 0.0;
      UB[a][i] = 0.0;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) RGN[i][j] = 0.0;
#pragma unroll
  for (int k = 0; k < 10; ++k) SUPG[k] = 0.0;

#pragma unroll
  for (int qp = 0; qp < 4; ++qp) {
    const double N0 = nq(qp, 0), N1 = nq(qp, 1), N2 = nq(qp, 2), N3 = nq(qp, 3);
    double u[3], aq[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      u[i] = N0 * ul[0][i] + N1 * ul[1][i] + N2 * ul[2][i] + N3 * ul[3][i];
      aq[i] = N0 * al[0][i] + N1 * al[1][i] + N2 * al[2][i] + N3 * al[3][i];
    }
    const double pq = N0 * pl[0] + N1 * pl[1] + N2 * pl[2] + N3 * pl[3];
    // strong residual (e3res.f90:775-789)
    double rL[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      rL[i] = (aq[i] + u[0] * gu[0][i] + u[1] * gu[1][i] + u[2] * gu[2][i]) * rho + gp[i] - divq[i];
    // tau, itau = 0 (e3stab.f90:108-133)
    const double uGu = u[0] * (gg[0] * u[0] + gg[3] * u[1] + gg[5] * u[2]) +
                       u[1] * (gg[3] * u[0] + gg[1] * u[1] + gg[4] * u[2]) +
                       u[2] * (gg[5] * u[0] + gg[4] * u[1] + gg[2] * u[2]);
    const double fact = sqrt((tau0 + uGu) + visc);
    const double tauC = rho * 0.125 * fact / trg * ff;
    const double tauM = 1.0 / fact;
    double tauBar = rL[0] * (gg[0] * rL[0] + gg[3] * rL[1] + gg[5] * rL[2]) +
                    rL[1] * (gg[3] * rL[0] + gg[1] * rL[1] + gg[4] * rL[2]) +
                    rL[2] * (gg[5] * rL[0] + gg[4] * rL[1] + gg[2] * rL[2]);
    tauBar = (tauBar != 0.0) ? tauM / sqrt(tauBar) : 0.0;
    double ub[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ub[i] = u[i] - tauM * rL[i] * rhoinv;
    // e3Res, iconvflow = 2 (e3res.f90:99-286)
    const double tmp = -pq + tauC * divu;
    double lg[3];  // L . grad u_i
#pragma unroll
    for (int i = 0; i < 3; ++i) lg[i] = tauBar * (rL[0] * gu[0][i] + rL[1] * gu[1][i] + rL[2] * gu[2][i]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double rna = (aq[i] + ub[0] * gu[0][i] + ub[1] * gu[1][i] + ub[2] * gu[2][i]) * rho;
      RNa[0][i] += N0 * rna;
      RNa[1][i] += N1 * rna;
      RNa[2][i] += N2 * rna;
      RNa[3][i] += N3 * rna;
      const double tl = tauM * rL[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double v = mu * (gu[i][j] + gu[j][i]) + tl * u[j] + lg[i] * rL[j];
        if (i == j) v += tmp;
        RGN[i][j] += v;
      }
      SU[i] += ub[i];
    }
    if (LHS) {
      // e3LHS, iconvflow = 2 (e3lhs.f90:124-129): convection by u-bar, and
      // the tauBar (L.g_a)(L.g_b) term beside SUPG's tauM rho (u.g_a)(u.g_b)
      TC += tauC;
      TM += tauM;
      const double tb = tauBar * rhoinv;
      double z[4], zl[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        z[a] = u[0] * g[a][0] + u[1] * g[a][1] + u[2] * g[a][2];
        zl[a] = rL[0] * g[a][0] + rL[1] * g[a][1] + rL[2] * g[a][2];
      }
      int k = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a; b < 4; ++b) SUPG[k++] += tauM * z[a] * z[b] + tb * zl[a] * zl[b];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        UB[0][i] += N0 * ub[i];
        UB[1][i] += N1 * ub[i];
        UB[2][i] += N2 * ub[i];
        UB[3][i] += N3 * ub[i];
      }
    }
  }

  double *r = LHS ? sh_rec + threadIdx.x * kE3Pad : rec + (size_t)e * kRec;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    r[R_A(a) + R_RL + 3] = W * (g[a][0] * SU[0] + g[a][1] * SU[1] + g[a][2] * SU[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      r[R_A(a) + R_RL + i] = -W * (RNa[a][i] + g[a][0] * RGN[i][0] + g[a][1] * RGN[i][1] + g[a][2] * RGN[i][2]);
  }
  if (!LHS) return;
  const double tlW = c.lhsFct * W;
  const double tsFct = c.lhmFct * W * rho;
  const double Mu = 4.0 * (tlW * mu);
  const double rt = rho * tlW;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      // convection sum_q N_a (ubar_q . g_b)
      const double conv = UB[a][0] * g[b][0] + UB[a][1] * g[b][1] + UB[a][2] * g[b][2];
      const int lo = a < b ? a : b, hi = a < b ? b : a;
      const int k = lo * 4 - (lo * (lo - 1)) / 2 + (hi - lo);  // index of (lo,hi) in SUPG
      const double gab = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
      r[R_A(a) + b] = tsFct * mm(a, b) + rt * conv + Mu * gab + rt * SUPG[k];
    }
  if (e < nel) {
    double2 *d = reinterpret_cast<double2 *>(esc + (size_t)e * kESc);
    d[0] = make_double2(tlW * TC, Mu);     // Tc, Mu
    d[1] = make_double2(tlW * TM / rho, tlW);  // Cm, tlW
  }
  __syncthreads();
  const int cnt = min(kE3Threads, nel - e0);
  double2 *dst = reinterpret_cast<double2 *>(rec + (size_t)e0 * kRec);
  for (int i = threadIdx.x; i < cnt * (kRec / 2); i += kE3Threads) {
    const double *src = sh_rec + (i / (kRec / 2)) * kE3Pad + 2 * (i % (kRec / 2));
    dst[i] = make_double2(src[0], src[1]);
  }
}

// ---------------------------------------------------------------- FP32 tangent
// The mixed-precision formation keeps the residual in FP64 (flux pass, e3
// residual, res_node) and forms the tangent in FP32 with FP32 accumulation.
// The iconvflow = 2 tangent needs the strong residual L (tauBar and u-bar,
// e3lhs.f90:124-129), so this kernel gathers the same node records as
// e3_elem (x, u, p, ac and the reconstructed flux q) and evaluates the metric,
// gradients, L, tau and u-bar at the 4 quadrature points and the element
// record S/g/scalars.  Elements whose aspect ratio exceeds a threshold are
// evaluated in FP64 and rounded to FP32 at the end (metric and tau on skewed
// elements are where FP32 loses most).
// FP32 record: 32 floats = one 128-byte line: g[a][j] at a*3+j, tlW, Tc, Mu,
// Cm at 12..15, S[a][0..3] at 16 + 4a.
constexpr int kRec32 = 32;
__host__ __device__ constexpr int R32_S(int a) { return 16 + 4 * a; }

// Longest edge over shortest altitude, normalised to 1 for the regular tet.
__device__ inline float tet_aspect(const double (&xd)[4][3]) {
  float x[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) x[a][i] = (float)xd[a][i];
  auto sub = [&](int a, int b, float (&d)[3]) {
    d[0] = x[a][0] - x[b][0];
    d[1] = x[a][1] - x[b][1];
    d[2] = x[a][2] - x[b][2];
  };
  auto cross2 = [](const float (&u)[3], const float (&v)[3]) {
    const float c0 = u[1] * v[2] - u[2] * v[1], c1 = u[2] * v[0] - u[0] * v[2], c2 = u[0] * v[1] - u[1] * v[0];
    return c0 * c0 + c1 * c1 + c2 * c2;
  };
  float e01[3], e02[3], e03[3], e12[3], e13[3];
  sub(1, 0, e01);
  sub(2, 0, e02);
  sub(3, 0, e03);
  sub(2, 1, e12);
  sub(3, 1, e13);
  float e23[3];
  sub(3, 2, e23);
  auto l2 = [](const float (&d)[3]) { return d[0] * d[0] + d[1] * d[1] + d[2] * d[2]; };
  const float L2 = fmaxf(fmaxf(fmaxf(l2(e01), l2(e02)), fmaxf(l2(e03), l2(e12))), fmaxf(l2(e13), l2(e23)));
  // squared doubled areas of the 4 faces
  const float A2 = fmaxf(fmaxf(cross2(e01, e02), cross2(e01, e03)), fmaxf(cross2(e02, e03), cross2(e12, e13)));
  float c[3];
  c[0] = e02[1] * e03[2] - e02[2] * e03[1];
  c[1] = e02[2] * e03[0] - e02[0] * e03[2];
  c[2] = e02[0] * e03[1] - e02[1] * e03[0];
  const float six_v = fabsf(e01[0] * c[0] + e01[1] * c[1] + e01[2] * c[2]);
  // h_min = 3V / A_max = six_v / sqrt(A2);  aspect = L_max / h_min / sqrt(3/2)
  if (!(six_v > 0.0f)) return 3.0e38f;
  return sqrtf(L2) * sqrtf(A2) / six_v * 0.8164965809277260f;
}

// T: arithmetic type of the tangent; nl: the element's four node records.
template <class T>
__device__ inline void tangent_body(const double (&nl)[4][kNRec], const FormConst &c, float *__restrict__ r) {
  T xl[4][3], ul[4][3], al[4][3], pl[4], ql[4][6];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xl[a][i] = (T)nl[a][N_X + i];
      ul[a][i] = (T)nl[a][N_U + i];
      al[a][i] = (T)nl[a][N_A + i];
    }
    pl[a] = (T)nl[a][N_P];
#pragma unroll
    for (int k = 0; k < 6; ++k) ql[a][k] = (T)nl[a][N_Q + k];
  }
  T dx[3][3], W, g[4][3], gg[6];
  p1_metric(xl, dx, W, g);
  gijd_tet(dx, gg);
  T gu[3][3], gp[3], divq[3] = {T(0), T(0), T(0)};  // gu[j][i] = du_i/dx_j
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
      gu[j][i] = g[0][j] * ul[0][i] + g[1][j] * ul[1][i] + g[2][j] * ul[2][i] + g[3][j] * ul[3][i];
    gp[j] = g[0][j] * pl[0] + g[1][j] * pl[1] + g[2][j] * pl[2] + g[3][j] * pl[3];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      divq[i] += g[a][0] * ql[a][qsym(i)] + g[a][1] * ql[a][qsym(i + 3)] + g[a][2] * ql[a][qsym(i + 6)];
  const T rho = (T)c.rho, mu = (T)c.mu, rhoinv = T(1) / rho;
  const T rnu = mu * rhoinv;
  const T dts = (T)(c.Dtgl * c.dtsfct);
  const T GG = gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2] + T(2) * (gg[3] * gg[3] + gg[4] * gg[4] + gg[5] * gg[5]);
  const T tau0 = (T(2) * dts) * (T(2) * dts);
  const T visc = T(36) * (rnu * rnu) * GG;
  const T trg = gg[0] + gg[1] + gg[2];
  const T ff = (T)(c.taucfct / c.dtsfct);
  T TC = T(0), TM = T(0), SUPG[10], UB[4][3];
#pragma unroll
  for (int k = 0; k < 10; ++k) SUPG[k] = T(0);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) UB[a][i] = T(0);
#pragma unroll
  for (int qp = 0; qp < 4; ++qp) {
    const T Nq[4] = {(T)nq(qp, 0), (T)nq(qp, 1), (T)nq(qp, 2), (T)nq(qp, 3)};
    T u[3], aq[3], rL[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      u[i] = Nq[0] * ul[0][i] + Nq[1] * ul[1][i] + Nq[2] * ul[2][i] + Nq[3] * ul[3][i];
      aq[i] = Nq[0] * al[0][i] + Nq[1] * al[1][i] + Nq[2] * al[2][i] + Nq[3] * al[3][i];
    }
    // strong residual (e3res.f90:775-789)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      rL[i] = (aq[i] + u[0] * gu[0][i] + u[1] * gu[1][i] + u[2] * gu[2][i]) * rho + gp[i] - divq[i];
    const T uGu = u[0] * (gg[0] * u[0] + gg[3] * u[1] + gg[5] * u[2]) +
                  u[1] * (gg[3] * u[0] + gg[1] * u[1] + gg[4] * u[2]) +
                  u[2] * (gg[5] * u[0] + gg[4] * u[1] + gg[2] * u[2]);
    const T fact = sqrt((tau0 + uGu) + visc);
    const T tauC = rho * T(0.125) * fact / trg * ff;
    const T tauM = T(1) / fact;
    T tauBar = rL[0] * (gg[0] * rL[0] + gg[3] * rL[1] + gg[5] * rL[2]) +
               rL[1] * (gg[3] * rL[0] + gg[1] * rL[1] + gg[4] * rL[2]) +
               rL[2] * (gg[5] * rL[0] + gg[4] * rL[1] + gg[2] * rL[2]);
    tauBar = (tauBar != T(0)) ? tauM / sqrt(tauBar) : T(0);
    TC += tauC;
    TM += tauM;
    const T tb = tauBar * rhoinv;
    T z[4], zl[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      z[a] = u[0] * g[a][0] + u[1] * g[a][1] + u[2] * g[a][2];
      zl[a] = rL[0] * g[a][0] + rL[1] * g[a][1] + rL[2] * g[a][2];
    }
    int k = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = a; b < 4; ++b) SUPG[k++] += tauM * z[a] * z[b] + tb * zl[a] * zl[b];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const T ub = u[i] - tauM * rL[i] * rhoinv;
#pragma unroll
      for (int a = 0; a < 4; ++a) UB[a][i] += Nq[a] * ub;
    }
  }
  const T tlW = (T)c.lhsFct * W;
  const T tsFct = (T)c.lhmFct * W * rho;
  const T Mu = T(4) * (tlW * mu);
  const T rt = rho * tlW;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const T conv = UB[a][0] * g[b][0] + UB[a][1] * g[b][1] + UB[a][2] * g[b][2];
      const int lo = a < b ? a : b, hi = a < b ? b : a;
      const int k = lo * 4 - (lo * (lo - 1)) / 2 + (hi - lo);
      const T gab = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
      r[R32_S(a) + b] = (float)(tsFct * (T)mm(a, b) + rt * conv + Mu * gab + rt * SUPG[k]);
    }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) r[R_G + a * 3 + j] = (float)g[a][j];
  r[R_TLW] = (float)tlW;
  r[R_TC] = (float)(tlW * TC);
  r[R_MU] = (float)Mu;
  r[R_CM] = (float)(tlW * TM / rho);
}

__global__ void __launch_bounds__(256) e3_tangent32(int nel, const int *__restrict__ ien,
                                                    const double *__restrict__ node, FormConst c0, Ens en,
                                                    float aspect_fp64, float *__restrict__ rec) {
  const int e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  FormConst c = c0;
  c.mu = ens_mu(en, e, c0.mu);
  double nl[4][kNRec];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double2 *rr = reinterpret_cast<const double2 *>(node + (size_t)ien[a * (size_t)nel + e] * kNRec);
#pragma unroll
    for (int k = 0; k < kNRec / 2; ++k) {
      const double2 v = rr[k];
      nl[a][2 * k] = v.x;
      nl[a][2 * k + 1] = v.y;
    }
  }
  double xl[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) xl[a][i] = nl[a][N_X + i];
  float r[kRec32];
  if (tet_aspect(xl) > aspect_fp64)
    tangent_body<double>(nl, c, r);
  else
    tangent_body<float>(nl, c, r);
  float4 *o = reinterpret_cast<float4 *>(rec + (size_t)e * kRec32);
#pragma unroll
  for (int k = 0; k < kRec32 / 4; ++k) o[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

// ---------------------------------------------------------------- assembly
// Output modes: 0 = reference layout lhsK(9,nnz)/lhsP(4,nnz);
// 1 = memLS planes V[16][nnz] (row-major 4x4 blocks of solfar's Val4).
struct AsmOut {
  double *lhsK, *lhsP;  // mode 0
  double *V;            // mode 1, FP64 matrix
  float *V32;           // mode 1, FP32 matrix
  size_t nnz;
};

// The 16 plane values of one block (row-major 4x4 of Val4): solfar.f90:216-221:
// Val4(1:3)=lhsK(1:3), Val4(5:7)=lhsK(4:6), Val4(9:11)=lhsK(7:9),
// Val4(13:15)=lhsP(1:3), Val4(16)=lhsP(4); Val4(4:12:4) = -lhsP(1:3) of the
// transposed block (:225-235).
template <class T>
__device__ inline void block_planes(const T (&K)[9], const T (&Gk)[3], T C, const T (&Gtr)[3], T (&v)[16]) {
  v[0] = K[0];
  v[1] = K[1];
  v[2] = K[2];
  v[3] = -Gtr[0];
  v[4] = K[3];
  v[5] = K[4];
  v[6] = K[5];
  v[7] = -Gtr[1];
  v[8] = K[6];
  v[9] = K[7];
  v[10] = K[8];
  v[11] = -Gtr[2];
  v[12] = Gk[0];
  v[13] = Gk[1];
  v[14] = Gk[2];
  v[15] = C;
}

// Write one block given K (PHASTA column-major order m = i + 3j), the
// pressure-gradient column G(k) = lhsP(1:3) of this block, C = lhsP(4), and
// the transposed block's lhsP(1:3) (for Val4's -G^T column).
template <int MODE, class T>
__device__ inline void put_block(const AsmOut &o, int k, const T (&K)[9], const T (&Gk)[3], T C,
                                 const T (&Gtr)[3]) {
  if (MODE == 0) {
    double *pk = o.lhsK + (size_t)k * 9;
#pragma unroll
    for (int m = 0; m < 9; ++m) pk[m] = K[m];
    double *pp = o.lhsP + (size_t)k * 4;
    pp[0] = Gk[0];
    pp[1] = Gk[1];
    pp[2] = Gk[2];
    pp[3] = C;
  } else {
    const size_t S = o.nnz;
    T *V = (sizeof(T) == 4 ? (T *)o.V32 : (T *)o.V) + k;
    T v[16];
    block_planes(K, Gk, C, Gtr, v);
#pragma unroll
    for (int m = 0; m < 16; ++m) V[m * S] = v[m];
  }
}

// Residual rows only (lhs = 0): res_i = sum of rl_a over the node's
// incidences (the scatter of asigmr.f90:116).  MODE 0 writes res(nshg,4) like
// ElmGMR; MODE 1 writes Res4(4,nshg), the transposed copy SolFlow hands to
// memLS (solfar.f90:206-210).
template <int MODE>
__device__ inline void put_res(double *res, int i, int nshg, const double (&rs)[4]) {
  if (MODE == 0) {
    const size_t N = nshg;
    res[i + 0 * N] = rs[0];
    res[i + 1 * N] = rs[1];
    res[i + 2 * N] = rs[2];
    res[i + 3 * N] = rs[3];
  } else {
    res[4 * (size_t)i + 0] = rs[0];
    res[4 * (size_t)i + 1] = rs[1];
    res[4 * (size_t)i + 2] = rs[2];
    res[4 * (size_t)i + 3] = rs[3];
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) res_node(int nshg, const int *__restrict__ inc_ptr,
                                                const int *__restrict__ inc, const double *__restrict__ rec,
                                                double *__restrict__ res) {
  const int gid = xcd_block() * blockDim.x + threadIdx.x;
  const int i = gid / kNodeLanes, lane = gid % kNodeLanes;
  double rs[4] = {0, 0, 0, 0};
  if (i < nshg)
    for (int p = inc_ptr[i] + lane; p < inc_ptr[i + 1]; p += kNodeLanes) {
      const int ea = inc[p];
      const double2 *r = reinterpret_cast<const double2 *>(rec + (size_t)(ea >> 2) * kRec + R_A(ea & 3) + R_RL);
      const double2 r01 = r[0], r23 = r[1];
      rs[0] += r01.x;
      rs[1] += r01.y;
      rs[2] += r23.x;
      rs[3] += r23.y;
    }
#pragma unroll
  for (int off = 1; off < kNodeLanes; off <<= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) rs[k] += __shfl_xor(rs[k], off, kNodeLanes);
  if (i < nshg && lane == 0) put_res<MODE>(res, i, nshg, rs);
}

// Row-parallel assembly.  Contribution of element incidence (staged record r,
// local nodes a, b) to the block (ien(e,a), ien(e,b)): K^{ii,jj} += delta S_ab
// + Tc g_a,ii g_b,jj + Mu g_a,jj g_b,ii, lhsP(1:3) += tlW sum_q N_a g_b, the
// transposed block's lhsP(1:3) += tlW sum_q N_b g_a, C += Cm g_a.g_b
// (e3lhs.f90:171-402).
template <class T>
struct BlockAcc {
  T K[9], Gij[3], Gji[3], C;
};
template <class T>
__device__ inline void block_zero(BlockAcc<T> &A) {
#pragma unroll
  for (int m = 0; m < 9; ++m) A.K[m] = T(0);
#pragma unroll
  for (int m = 0; m < 3; ++m) A.Gij[m] = A.Gji[m] = T(0);
  A.C = T(0);
}
__device__ inline double sum_n_rt(int a) {
  return a == 0 ? sum_n(0) : a == 1 ? sum_n(1) : a == 2 ? sum_n(2) : sum_n(3);
}
// Staged incidence: g, tlW, Tc, Mu, Cm at 0..15 and node a's S row at 16..19
// (the residual row goes straight to the lane that sums it).
constexpr int kStage = 20;
constexpr int ST_S = 16;
template <class T>
__device__ inline void block_add(BlockAcc<T> &A, const T *r, int a, int b) {
  const T ga[3] = {r[R_G + a * 3 + 0], r[R_G + a * 3 + 1], r[R_G + a * 3 + 2]};
  const T gb[3] = {r[R_G + b * 3 + 0], r[R_G + b * 3 + 1], r[R_G + b * 3 + 2]};
  const T Sab = r[ST_S + b], Tc = r[R_TC], Mu = r[R_MU], tl = r[R_TLW], Cm = r[R_CM];
#pragma unroll
  for (int jj = 0; jj < 3; ++jj)
#pragma unroll
    for (int ii = 0; ii < 3; ++ii)
      A.K[ii + 3 * jj] += (ii == jj ? Sab : T(0)) + (Tc * ga[ii] * gb[jj] + Mu * ga[jj] * gb[ii]);
  const T ta = tl * T(sum_n_rt(a)), tb = tl * T(sum_n_rt(b));
#pragma unroll
  for (int ii = 0; ii < 3; ++ii) {
    A.Gij[ii] += ta * gb[ii];
    A.Gji[ii] += tb * ga[ii];
  }
  A.C += Cm * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
}

// Diagonal contribution (b = a): the transposed block is the block itself,
// so only K, lhsP(1:3) and C are accumulated.  K^{ii,jj} = delta S_aa +
// Tc g_a,ii g_a,jj + Mu g_a,jj g_a,ii is symmetric: its upper triangle (ii <=
// jj, each entry formed as the full-block loop forms it) is kept, in the
// order 00 01 02 11 12 22, and mirrored when the block is written.
template <class T>
struct DiagAcc {
  T K[6], G[3], C;
};
__host__ __device__ constexpr int sym3(int ii, int jj) {
  return ii <= jj ? ii * 3 - (ii * (ii - 1)) / 2 + (jj - ii) : jj * 3 - (jj * (jj - 1)) / 2 + (ii - jj);
}
template <class T>
__device__ inline void diag_zero(DiagAcc<T> &A) {
#pragma unroll
  for (int m = 0; m < 6; ++m) A.K[m] = T(0);
  A.G[0] = A.G[1] = A.G[2] = A.C = T(0);
}
template <class T>
__device__ inline void diag_add(DiagAcc<T> &A, const T *r, int a) {
  const T ga[3] = {r[R_G + a * 3 + 0], r[R_G + a * 3 + 1], r[R_G + a * 3 + 2]};
  const T Saa = r[ST_S + a], Tc = r[R_TC], Mu = r[R_MU], tl = r[R_TLW], Cm = r[R_CM];
#pragma unroll
  for (int jj = 0; jj < 3; ++jj)
#pragma unroll
    for (int ii = 0; ii <= jj; ++ii)
      A.K[sym3(ii, jj)] += (ii == jj ? Saa : T(0)) + (Tc * ga[ii] * ga[jj] + Mu * ga[jj] * ga[ii]);
  const T ta = tl * T(sum_n_rt(a));
#pragma unroll
  for (int ii = 0; ii < 3; ++ii) A.G[ii] += ta * ga[ii];
  A.C += Cm * (ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2]);
}
// The lanes' diagonal partial sums combined by a fixed shuffle tree over
// the kRowLanes lanes of a row.
template <class T>
__device__ inline void diag_reduce(DiagAcc<T> &D) {
#pragma unroll
  for (int off = kRowLanes / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int m = 0; m < 6; ++m) D.K[m] += __shfl_xor(D.K[m], off, kRowLanes);
#pragma unroll
    for (int m = 0; m < 3; ++m) D.G[m] += __shfl_xor(D.G[m], off, kRowLanes);
    D.C += __shfl_xor(D.C, off, kRowLanes);
  }
}
template <class T>
__device__ inline void diag_full(const DiagAcc<T> &D, T (&K)[9]) {
#pragma unroll
  for (int jj = 0; jj < 3; ++jj)
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) K[ii + 3 * jj] = D.K[sym3(ii, jj)];
}

// Orders LDS accesses of the lanes of one wavefront (the lanes of a CSR row
// always share a wavefront).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A lane's contribution-list entries (16 bits each), eight at a time: two
// 64-bit words filled by four independent dword loads, consumed from the low
// end.  The list of one block is contiguous in lst and sorted by incidence,
// so the loads of a whole block are issued at once instead of one dependent
// load per contribution.  Only dwords holding an entry of [q, q1) are read.
struct LstBuf {
  unsigned long long w0, w1;
  int n;     // entries left in the buffer
  int skip;  // 1: the first 16 bits precede q (q odd); dropped at first use
};
__device__ inline void lst_fill(LstBuf &b, const unsigned short *__restrict__ lst, int q, int q1) {
  const unsigned *lw = reinterpret_cast<const unsigned *>(lst);
  const int d = q >> 1, dend = (q1 + 1) >> 1;  // dwords [d, dend) hold the entries
  unsigned v[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) v[m] = d + m < dend ? lw[d + m] : 0u;
  b.w0 = (unsigned long long)v[0] | (unsigned long long)v[1] << 32;
  b.w1 = (unsigned long long)v[2] | (unsigned long long)v[3] << 32;
  b.skip = q & 1;
  b.n = min(q1 - q, 8 - (q & 1));
}
__device__ inline void lst_shift(LstBuf &b) {
  b.w0 = b.w0 >> 16 | b.w1 << 48;
  b.w1 >>= 16;
}
// the next entry (the buffer holds at least one)
__device__ inline unsigned lst_front(LstBuf &b) {
  if (b.skip) {
    lst_shift(b);
    b.skip = 0;
  }
  return (unsigned)(b.w0 & 0xffffu);
}
__device__ inline void lst_pop(LstBuf &b) {
  lst_shift(b);
  --b.n;
}

// Per row of the schedule: row, first CSR position, slots, diagonal slot
// (int4), first incidence and incidence count (int4, two unused), built once
// with the schedule so that a row's assembly starts from one load.
__global__ void slot_table_kernel(int n, const int *__restrict__ order, const int *__restrict__ rowptr,
                                  const int *__restrict__ diagk, const int *__restrict__ inc_ptr,
                                  int4 *__restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int row = order[i];
  const int k0 = rowptr[row], p0 = inc_ptr[row];
  tab[2 * i] = make_int4(row, k0, rowptr[row + 1] - k0, diagk[row] - k0);
  tab[2 * i + 1] = make_int4(p0, inc_ptr[row + 1] - p0, 0, 0);
}

// Row-parallel assembly, kRowLanes lanes per CSR row; rows are taken in the
// handle's schedule order (neighbouring rows together, so that the element
// records and packed nodes they share are read from L2).  The row's element
// incidences are taken in chunks of CH, in element order; lane t of a chunk
// takes incidence t: it loads node a's S row and the element scalars,
// rebuilds the element's shape gradients from its packed nodes (ienp: the
// element's four nodes, stored per incidence so that a row reads them as one
// contiguous run) and stages g, the scalars and the S row in LDS.  Then:
//  - lane s accumulates off-diagonal block s (and s + 16, ... for long rows)
//    from its contribution list (lst: local incidence << 4 | a << 2 | b, in
//    element order, built once from the sparseloc positions): fillsparseI's
//    accumulation (fillsparse.f90:29-63) for that block;
//  - lane t < CH adds incidence t's diagonal contribution (b = a) and
//    residual row (asigmr.f90:116); the lanes' partial sums are combined by a
//    fixed shuffle tree.
// Every block of the row is written, so blocks no element touches come out
// zero (the reference zeroes lhsK/lhsP first, elmgmr.f90:127-130).
constexpr int kRowsPerBlock = 256 / kRowLanes;
constexpr int kAsmChunk = 16;  // incidences staged per chunk (even, dividing kRowLanes)
template <int MODE, int CH = kAsmChunk>
__global__ void __launch_bounds__(256) asm_row(int nshg, const int4 *__restrict__ tab, const int *__restrict__ inc,
                                               const int4 *__restrict__ ienp, const double *__restrict__ xu,
                                               const int *__restrict__ lst_ptr, const unsigned short *__restrict__ lst,
                                               const double *__restrict__ rec, const double *__restrict__ esc,
                                               double *__restrict__ res, AsmOut o, unsigned xc, int runs) {
  static_assert(CH % 2 == 0 && kRowLanes % CH == 0, "chunk of 2, 4, 8 or 16 incidences");
  __shared__ double st_all[kRowsPerBlock][CH][kStage];
  const int gid = xcd_block(xc) * blockDim.x + threadIdx.x;
  const int slot = gid / kRowLanes, lane = gid % kRowLanes;
  const bool live = slot < nshg;
  double(*st)[kStage] = st_all[threadIdx.x / kRowLanes];
  int row = 0, k0 = 0, nslot = 0, sdiag = -1, p0 = 0, ninc = 0;
  if (live) {
    const int4 t0 = tab[2 * slot], t1 = tab[2 * slot + 1];
    row = t0.x;
    k0 = t0.y;
    nslot = t0.z;
    sdiag = t0.w;
    p0 = t1.x;
    ninc = t1.y;
  }
  const int nchunk = (ninc + CH - 1) / CH;
  const int ngroup = (nslot + kRowLanes - 1) / kRowLanes;
  // Staged plane stores (planes, chunks of 16): when the wave's four rows
  // are consecutive rows (runs: the schedule is the numbering) of at most
  // 16 blocks each, their blocks form one range of every plane; the wave
  // gathers them, the diagonal blocks included, in its staging LDS after
  // the rows are summed and stores each plane's range at once, instead of
  // the off-diagonal blocks and, later, the diagonal ones into the holes
  // they left.
  const bool wstg = MODE == 1 && CH == kRowLanes && runs && __all(nslot <= kRowLanes);
  DiagAcc<double> D;
  diag_zero(D);
  double rs[4] = {0, 0, 0, 0};
  BlockAcc<double> A;
  for (int g = 0; g < ngroup; ++g) {
    const int s = lane + g * kRowLanes;
    const bool own = s < nslot && s != sdiag;
    block_zero(A);
    // the block's list range and the first chunk's incidences, loaded
    // together (lanes without a block read lst_ptr[0..1], an in-bounds pair;
    // lanes without an incidence read incidence 0, always present, and stage
    // nothing of it)
    const int kq = own ? k0 + s : 0;
    int q = lst_ptr[kq];
    const int q1r = lst_ptr[kq + 1];
    int nt = min(CH, ninc);
    int pi = lane < nt ? p0 + lane : 0;
    int ea = inc[pi];
    int4 n4 = ienp[pi];
    const int q1 = own ? q1r : q;
    LstBuf lb;
    lst_fill(lb, lst, q, q1);
    for (int c = 0; c < nchunk; ++c) {
      const int tb = c * CH;
      // incidence t of the row is held by lane t % 16 whatever the chunk
      // size, so that every chunk size sums the diagonal block and the
      // residual row in the same order: chunk c occupies lanes lo..lo+CH-1
      const int lo = CH == kRowLanes ? 0 : tb % kRowLanes;
      const int lon = CH == kRowLanes ? 0 : (tb + CH) % kRowLanes;
      const bool mine = lane >= lo && lane < lo + CH;  // the lane holds a staging slot of this chunk
      const int t = lane - lo;
      // the next chunk's incidences, in flight while this chunk is staged
      const int ntn = min(CH, ninc - tb - CH);
      const int pn = lane >= lon && lane - lon < ntn ? p0 + tb + CH + lane - lon : 0;
      const int ean = inc[pn];
      const int4 n4n = ienp[pn];
      // node a's S row, the element scalars and the nodes' coordinates
      const int a = ea & 3;
      const double *ra = rec + (size_t)(ea >> 2) * kRec + R_A(a);
      const double2 *sc = reinterpret_cast<const double2 *>(esc + (size_t)(ea >> 2) * kESc);
      const double2 s01 = *reinterpret_cast<const double2 *>(ra);
      const double2 s23 = *reinterpret_cast<const double2 *>(ra + 2);
      const double2 sc0 = sc[0], sc1 = sc[1];  // (Tc, Mu), (Cm, tlW)
      double xl[4][3];
      gather_x(xu, n4, xl);
      double dx[3][3], W, gl[4][3];
      p1_metric(xl, dx, W, gl);
      wave_lds_sync();  // the previous chunk's readers are done
      if (mine) {
        double *d = st[t];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int j = 0; j < 3; ++j) d[R_G + 3 * b + j] = gl[b][j];
        d[R_TC] = sc0.x;
        d[R_MU] = sc0.y;
        d[R_CM] = sc1.x;
        d[R_TLW] = sc1.y;
        d[ST_S + 0] = s01.x;
        d[ST_S + 1] = s01.y;
        d[ST_S + 2] = s23.x;
        d[ST_S + 3] = s23.y;
      }
      wave_lds_sync();
      // the incidence's own lane: diagonal contribution, and its residual
      // row, loaded here and added after the off-diagonal blocks so that the
      // load is not held in registers while the chunk is staged
      const bool dl = g == 0 && mine && t < nt;
      double2 rl01 = make_double2(0.0, 0.0), rl23 = rl01;
      if (dl) {
        rl01 = *reinterpret_cast<const double2 *>(ra + R_RL);
        rl23 = *reinterpret_cast<const double2 *>(ra + R_RL + 2);
        diag_add(D, st[t], a);
      }
      while (q < q1) {
        if (lb.n == 0) lst_fill(lb, lst, q, q1);  // blocks with more than 8 contributions
        const unsigned v = lst_front(lb);
        const int tt = (int)(v >> 4) - tb;
        if (tt >= CH) break;
        block_add(A, st[tt], (int)((v >> 2) & 3u), (int)(v & 3u));
        lst_pop(lb);
        ++q;
      }
      if (dl) {
        rs[0] += rl01.x;
        rs[1] += rl01.y;
        rs[2] += rl23.x;
        rs[3] += rl23.y;
      }
      nt = ntn;
      ea = ean;
      n4 = n4n;
    }
    if (own && !wstg) put_block<MODE>(o, k0 + s, A.K, A.Gij, A.C, A.Gji);
  }
  diag_reduce(D);
#pragma unroll
  for (int off = kRowLanes / 2; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) rs[k] += __shfl_xor(rs[k], off, kRowLanes);
  if constexpr (MODE == 1 && CH == kRowLanes) {
    if (wstg) {
      // the wave's range [kA, kB) of every plane (its live rows are
      // consecutive, so their blocks are)
      int kA = live ? k0 : 0x7fffffff, kB = live ? k0 + nslot : 0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        kA = min(kA, __shfl_xor(kA, off, 64));
        kB = max(kB, __shfl_xor(kB, off, 64));
      }
      const int nb = kB - kA;  // at most 4 rows x 16 blocks
      double *ob = &st_all[threadIdx.x / 64 * (64 / kRowLanes)][0][0];  // the wave's rows' staging: [plane][block - kA]
      wave_lds_sync();  // the last chunk's readers are done
      double v[16];
      if (lane < nslot && lane != sdiag) {
        block_planes(A.K, A.Gij, A.C, A.Gji, v);
#pragma unroll
        for (int m = 0; m < 16; ++m) ob[m * nb + k0 + lane - kA] = v[m];
      }
      if (live && lane == 0) {
        double K[9];
        diag_full(D, K);
        block_planes(K, D.G, D.C, D.G, v);
#pragma unroll
        for (int m = 0; m < 16; ++m) ob[m * nb + k0 + sdiag - kA] = v[m];
        put_res<MODE>(res, row, nshg, rs);
      }
      wave_lds_sync();
      const int wl = threadIdx.x % 64;
      if (wl < nb)
#pragma unroll
        for (int m = 0; m < 16; ++m) o.V[m * o.nnz + kA + wl] = ob[m * nb + wl];
      return;
    }
  }
  if (live && lane == 0) {
    double K[9];
    diag_full(D, K);
    put_block<MODE>(o, k0 + sdiag, K, D.G, D.C, D.G);
    put_res<MODE>(res, row, nshg, rs);
  }
}

// FP32 tangent assembly (no residual rows): the structure of asm_row with
// FP32 records (one line per element: g, scalars, S), FP32 accumulation and
// FP32 planes (MODE 1) or FP64 lhsK/lhsP (MODE 0).
template <int MODE>
__global__ void __launch_bounds__(256) asm_row32(int nshg, const int4 *__restrict__ tab,
                                                 const int *__restrict__ inc, const int *__restrict__ lst_ptr,
                                                 const unsigned short *__restrict__ lst,
                                                 const float *__restrict__ rec, AsmOut o, unsigned xc) {
  __shared__ float st_all[kRowsPerBlock][kRowLanes][kStage];
  const int gid = xcd_block(xc) * blockDim.x + threadIdx.x;
  const int slot = gid / kRowLanes, lane = gid % kRowLanes;
  const bool live = slot < nshg;
  float(*st)[kStage] = st_all[threadIdx.x / kRowLanes];
  int k0 = 0, nslot = 0, sdiag = -1, p0 = 0, ninc = 0;
  if (live) {
    const int4 t0 = tab[2 * slot], t1 = tab[2 * slot + 1];
    k0 = t0.y;
    nslot = t0.z;
    sdiag = t0.w;
    p0 = t1.x;
    ninc = t1.y;
  }
  const int nchunk = (ninc + kRowLanes - 1) / kRowLanes;
  const int ngroup = (nslot + kRowLanes - 1) / kRowLanes;
  DiagAcc<float> D;
  diag_zero(D);
  for (int g = 0; g < ngroup; ++g) {
    const int s = lane + g * kRowLanes;
    const bool own = s < nslot && s != sdiag;
    BlockAcc<float> A;
    block_zero(A);
    const int kq = own ? k0 + s : 0;
    int q = lst_ptr[kq];
    const int q1r = lst_ptr[kq + 1];
    int nt = min(kRowLanes, ninc);
    int ea = lane < nt ? inc[p0 + lane] : 0;
    const int q1 = own ? q1r : q;
    LstBuf lb;
    lst_fill(lb, lst, q, q1);
    for (int c = 0; c < nchunk; ++c) {
      const int tb = c * kRowLanes;
      const int ntn = min(kRowLanes, ninc - tb - kRowLanes);
      const int ean = lane < ntn ? inc[p0 + tb + kRowLanes + lane] : 0;
      wave_lds_sync();
      // line 0 of the record (4 pieces of 16 bytes) and node a's S row,
      // unguarded as in asm_row (past the chunk's end: incidence 0)
      float4 pc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int t = (j * kRowLanes + lane) >> 2, part = lane & 3;
        const int et = __shfl(ea, t, kRowLanes);
        pc[j] = *reinterpret_cast<const float4 *>(rec + (size_t)(et >> 2) * kRec32 + 4 * part);
      }
      const float4 sv = *reinterpret_cast<const float4 *>(rec + (size_t)(ea >> 2) * kRec32 + R32_S(ea & 3));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int t = (j * kRowLanes + lane) >> 2, part = lane & 3;
        st[t][4 * part + 0] = pc[j].x;
        st[t][4 * part + 1] = pc[j].y;
        st[t][4 * part + 2] = pc[j].z;
        st[t][4 * part + 3] = pc[j].w;
      }
      st[lane][ST_S + 0] = sv.x;
      st[lane][ST_S + 1] = sv.y;
      st[lane][ST_S + 2] = sv.z;
      st[lane][ST_S + 3] = sv.w;
      wave_lds_sync();
      if (g == 0 && lane < nt) diag_add(D, st[lane], ea & 3);
      while (q < q1) {
        if (lb.n == 0) lst_fill(lb, lst, q, q1);
        const unsigned v = lst_front(lb);
        const int t = (int)(v >> 4) - tb;
        if (t >= kRowLanes) break;
        block_add(A, st[t], (int)((v >> 2) & 3u), (int)(v & 3u));
        lst_pop(lb);
        ++q;
      }
      nt = ntn;
      ea = ean;
    }
    if (own) put_block<MODE>(o, k0 + s, A.K, A.Gij, A.C, A.Gji);
  }
  diag_reduce(D);
  if (live && lane == 0) {
    float K[9];
    diag_full(D, K);
    put_block<MODE>(o, k0 + sdiag, K, D.G, D.C, D.G);
  }
}

// bc3Res for codes 0/7 on the velocity bits, bit 2 (pressure) and bit 11
// (bc3res.f90:34-80).
// bc3Res (bc3res.f90:34-120) without periodic nodes: the pressure row of
// bit 2; the velocity rows of 1- and 2-component codes projected with
// BC(:,4:6) (bc(nshg, ndofbc); null when no node has such a code), of
// 3-component codes and of bit 11 zeroed.
__global__ void bc3res_kernel(int nshg, const int *__restrict__ ibc, const double *__restrict__ bc,
                              double *__restrict__ res, int res4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshg) return;
  const int c = ibc[i];
  const size_t N = nshg;
  // component k of node i: res(i,k) or Res4(k,i)
  const size_t s = res4 ? 1 : N, b = res4 ? 4 * (size_t)i : (size_t)i;
  if ((c >> 2) & 1) res[b + 3 * s] = 0.0;
  const int vb = (c >> 3) & 7;
  if (vb != 0 && vb != 7) {
    const double b4 = bc[i + 3 * N], b5 = bc[i + 4 * N], b6 = bc[i + 5 * N];
    double r1 = res[b], r2 = res[b + s], r3 = res[b + 2 * s];
    switch (vb) {
      case 1:
        r2 = r2 - b4 * r1;
        r3 = r3 - b5 * r1;
        r1 = 0.0;
        break;
      case 2:
        r1 = r1 - b4 * r2;
        r3 = r3 - b5 * r2;
        r2 = 0.0;
        break;
      case 3:
        r3 = r3 - b4 * r1 - b6 * r2;
        r1 = r2 = 0.0;
        break;
      case 4:
        r1 = r1 - b4 * r3;
        r2 = r2 - b5 * r3;
        r3 = 0.0;
        break;
      case 5:
        r2 = r2 - b4 * r1 - b6 * r3;
        r1 = r3 = 0.0;
        break;
      default:  // 6
        r1 = r1 - b4 * r2 - b6 * r3;
        r2 = r3 = 0.0;
        break;
    }
    res[b] = r1;
    res[b + s] = r2;
    res[b + 2 * s] = r3;
  }
  if (vb == 7 || ((c >> 11) & 1)) {
    res[b] = 0.0;
    res[b + s] = 0.0;
    res[b + 2 * s] = 0.0;
  }
}

// bc3lhs (bc3lhs.f90:41-690) after the assembly.  For P1 elements each of
// the reference's per-element operations is linear in the element block:
// the column operations of node j (M -> M T_j) and the row operations of
// node i on block (i, j) summed over the elements are those operations on
// the assembled block, and the constrained diagonal entries it sets to one
// per element sum to the node's element count.  The operations of a row and
// a column commute, except the code-6 row operations (which the setup
// rejects where the order would matter).  Entries e = 0..8 of a block are
// the reference's 1..9 (1 2 3 / 4 5 6 / 7 8 9, lhsK(1:9)).  Each operation:
// entry t -= BC(4 + b1) K(s1) [- BC(4 + b2) K(s2)].
struct BclOp {
  signed char t, s1, s2, b1, b2;  // s2 < 0: none
};
struct BclCode {
  BclOp col[6], row[6];
  signed char ncol, nrow, one0, one1;  // one1 < 0: one entry only
  unsigned short colz, rowz;           // zeroed entries (bit e)
};
__constant__ BclCode kBclCodes[7] = {
    {{}, {}, 0, 0, -1, -1, 0, 0},
    // 1: x1 (bc3lhs.f90:56-200)
    {{{1, 0, -1, 0, 0}, {4, 3, -1, 0, 0}, {7, 6, -1, 0, 0}, {2, 0, -1, 1, 0}, {5, 3, -1, 1, 0}, {8, 6, -1, 1, 0}},
     {{3, 0, -1, 0, 0}, {4, 1, -1, 0, 0}, {5, 2, -1, 0, 0}, {6, 0, -1, 1, 0}, {7, 1, -1, 1, 0}, {8, 2, -1, 1, 0}},
     6, 6, 0, -1, 0x049, 0x007},
    // 2: x2 (:202-297)
    {{{0, 1, -1, 0, 0}, {3, 4, -1, 0, 0}, {6, 7, -1, 0, 0}, {2, 1, -1, 1, 0}, {5, 4, -1, 1, 0}, {8, 7, -1, 1, 0}},
     {{0, 3, -1, 0, 0}, {1, 4, -1, 0, 0}, {2, 5, -1, 0, 0}, {6, 3, -1, 1, 0}, {7, 4, -1, 1, 0}, {8, 5, -1, 1, 0}},
     6, 6, 4, -1, 0x092, 0x038},
    // 3: x1 and x2 (:394-501)
    {{{2, 0, 1, 0, 2}, {5, 3, 4, 0, 2}, {8, 6, 7, 0, 2}},
     {{6, 0, 3, 0, 2}, {7, 1, 4, 0, 2}, {8, 2, 5, 0, 2}},
     3, 3, 0, 4, 0x0db, 0x03f},
    // 4: x3 (:299-392)
    {{{0, 2, -1, 0, 0}, {3, 5, -1, 0, 0}, {6, 8, -1, 0, 0}, {1, 2, -1, 1, 0}, {4, 5, -1, 1, 0}, {7, 8, -1, 1, 0}},
     {{0, 6, -1, 0, 0}, {1, 7, -1, 0, 0}, {2, 8, -1, 0, 0}, {3, 6, -1, 1, 0}, {4, 7, -1, 1, 0}, {5, 8, -1, 1, 0}},
     6, 6, 8, -1, 0x124, 0x1c0},
    // 5: x1 and x3 (:503-589)
    {{{1, 0, 2, 0, 2}, {4, 3, 5, 0, 2}, {7, 6, 8, 0, 2}},
     {{3, 0, 6, 0, 2}, {4, 1, 7, 0, 2}, {5, 2, 8, 0, 2}},
     3, 3, 0, 8, 0x16d, 0x1c7},
    // 6: x2 and x3 (:591-690), the row operations as the reference writes them
    {{{0, 1, 2, 0, 2}, {3, 4, 5, 0, 2}, {6, 7, 8, 0, 2}},
     {{6, 0, -1, 0, 0}, {7, 1, -1, 0, 0}, {8, 2, 5, 0, 2}},
     3, 3, 4, 8, 0x1b6, 0x03f},
};
template <class T>
__device__ inline void bcl_apply(T (&K)[9], const BclOp *ops, int n, const double (&bc)[3]) {
  for (int o = 0; o < n; ++o) {
    const BclOp op = ops[o];
    T v = K[op.t] - (T)bc[op.b1] * K[op.s1];
    if (op.s2 >= 0) v = v - (T)bc[op.b2] * K[op.s2];
    K[op.t] = v;
  }
}
// One thread per touched block {k, partial index of the row node, of the
// column node (-1: none), diagonal}.  MODE 0: lhsK(9, nnz) (double); MODE 1:
// the memLS planes of T (lhsK(1:9) in planes 0-2, 4-6, 8-10, put_block's).
template <int MODE, class T>
__global__ void bc3lhs_kernel(int n, const int4 *__restrict__ blk, const BclNode *__restrict__ nodes, T *out,
                              size_t nnz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int4 b = blk[t];
  const size_t k = (size_t)b.x;
  T K[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) K[e] = MODE == 0 ? out[9 * k + e] : out[(size_t)(e + e / 3) * nnz + k];
  if (b.z >= 0) {
    const BclNode nd = nodes[b.z];
    const BclCode &c = kBclCodes[nd.code];
    bcl_apply(K, c.col, c.ncol, nd.b);
#pragma unroll
    for (int e = 0; e < 9; ++e)
      if ((c.colz >> e) & 1) K[e] = T(0);
  }
  if (b.y >= 0) {
    const BclNode nd = nodes[b.y];
    const BclCode &c = kBclCodes[nd.code];
    bcl_apply(K, c.row, c.nrow, nd.b);
#pragma unroll
    for (int e = 0; e < 9; ++e)
      if ((c.rowz >> e) & 1) K[e] = T(0);
    if (b.w) {
      K[c.one0] = (T)nd.count;
      if (c.one1 >= 0) K[c.one1] = (T)nd.count;
    }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    if (MODE == 0)
      out[9 * k + e] = K[e];
    else
      out[(size_t)(e + e / 3) * nnz + k] = K[e];
  }
}

}  // namespace


namespace {

int validate_params(const crimson_form_params *p) {
  if (!p) {
    crimson::set_error("null parameters");
    return CRIMSON_EARG;
  }
  if (p->itau != 0 || p->iconvflow != 2 || p->idiff != 1 || p->ipord != 1 || p->flmpl != 0.0 ||
      p->flmpr != 0.0) {
    crimson::set_error(
        "unsupported flags (itau=%d iconvflow=%d idiff=%d ipord=%d flmpl=%g flmpr=%g); supported: "
        "itau=0 iconvflow=2 idiff=1 ipord=1 flmpl=flmpr=0",
        p->itau, p->iconvflow, p->idiff, p->ipord, p->flmpl, p->flmpr);
    return CRIMSON_EUNSUPPORTED;
  }
  if (p->ideformwall != 0 && p->ideformwall != 1) {
    crimson::set_error("ideformwall = %d: 0 (rigid walls) or 1 (deformable walls)", p->ideformwall);
    return CRIMSON_EARG;
  }
  if (p->precision != 64 && p->precision != 32) {
    crimson::set_error("formation precision %d not available (64: FP64; 32: FP64 residual, FP32 tangent)",
                       p->precision);
    return CRIMSON_EUNSUPPORTED;
  }
  if (p->precision == 32 && !(p->aspect_fp64 >= 1.0)) {
    crimson::set_error("aspect_fp64 = %g: the FP64 tau threshold must be >= 1", p->aspect_fp64);
    return CRIMSON_EARG;
  }
  if (!(p->rho > 0) || !(p->mu >= 0) || !(p->Delt > 0) || !(p->dtsfct > 0)) {
    crimson::set_error("non-physical parameters rho=%g mu=%g Delt=%g dtsfct=%g", p->rho, p->mu, p->Delt,
                       p->dtsfct);
    return CRIMSON_EARG;
  }
  return CRIMSON_OK;
}


// position of column `col` in 0-based CSR row `row`, or -1
inline int find_col(const std::vector<int> &rp, const std::vector<int> &cols, int row, int col) {
  auto b = cols.begin() + rp[row], e = cols.begin() + rp[row + 1];
  auto it = std::lower_bound(b, e, col);
  if (it == e || *it != col) return -1;
  return (int)(it - cols.begin());
}

// Row schedule of the assembly, built on the first call and kept: the
// caller's node numbering when setup found it local (see kLocalSpan), else
// the nodes in Morton (Z-curve) order of the first call's coordinates.  Only
// the order in which rows are processed depends on it, never a result.  On
// the 64x64x320 cylinder (structured numbering, elements spanning one
// z-layer of about 3,400 nodes) the numbering's order assembles in 1.58 ms
// against Morton order's 1.85: each row's plane segments then join their
// neighbours' into whole lines, while the record lines it reads twice stay
// in the last-level cache.
constexpr int kLocalSpan = 32768;
int ensure_order(crimson_form *f, const double *d_x) {
  if (f->d_order.p) return CRIMSON_OK;
  const int n = f->nshg;
  std::vector<int> order(n);
  if (f->sched == kSchedNumbering) {
    for (int i = 0; i < n; ++i) order[i] = i;
  } else {
    const size_t np = f->numnp;
    std::vector<double> x(np * 3);
    CR_HIP(hipMemcpyAsync(x.data(), d_x, x.size() * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    CR_HIP(hipStreamSynchronize(f->stream));
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = hi[k] = x[k * np];
      for (int i = 0; i < n; ++i) {
        lo[k] = std::min(lo[k], x[i + k * np]);
        hi[k] = std::max(hi[k], x[i + k * np]);
      }
    }
    auto spread = [](unsigned long long v) {  // 21 bits -> every third bit
      v &= 0x1fffffULL;
      v = (v | v << 32) & 0x1f00000000ffffULL;
      v = (v | v << 16) & 0x1f0000ff0000ffULL;
      v = (v | v << 8) & 0x100f00f00f00f00fULL;
      v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
      v = (v | v << 2) & 0x1249249249249249ULL;
      return v;
    };
    auto morton = [&](const double (&p)[3]) {
      unsigned long long m = 0;
      for (int k = 0; k < 3; ++k) {
        const double span = hi[k] - lo[k];
        const double t = span > 0 ? (p[k] - lo[k]) / span : 0.0;
        m |= spread((unsigned long long)(t * 2097151.0)) << k;
      }
      return m;
    };
    // Morton: every row by its node's code; tiles: runs of kSchedTile
    // consecutive rows (whose plane segments are contiguous) by the code of
    // their centroid, rows in order within a run
    const int unit = f->sched == kSchedMorton ? 1 : kSchedTile;
    const int nu = (n + unit - 1) / unit;
    std::vector<std::pair<unsigned long long, int>> key(nu);
    for (int u = 0; u < nu; ++u) {
      const int i0 = u * unit, i1 = std::min(n, i0 + unit);
      double c[3] = {0, 0, 0};
      for (int i = i0; i < i1; ++i)
        for (int k = 0; k < 3; ++k) c[k] += x[i + k * np];
      for (int k = 0; k < 3; ++k) c[k] /= (i1 - i0);
      key[u] = {morton(c), u};
    }
    std::sort(key.begin(), key.end());
    int p = 0;
    for (int u = 0; u < nu; ++u)
      for (int i = key[u].second * unit; i < std::min(n, (key[u].second + 1) * unit); ++i) order[p++] = i;
  }
  CR_TRY(f->d_order.upload(order.data(), n, f->stream));
  CR_TRY(f->d_slot.alloc((size_t)2 * n));
  slot_table_kernel<<<crimson::grid_for(n, 256), 256, 0, f->stream>>>(n, f->d_order.p, f->d_rowptr.p, f->d_diagk.p,
                                                                      f->d_inc_ptr.p, f->d_slot.p);
  CR_HIP(hipGetLastError());
  CR_HIP(hipStreamSynchronize(f->stream));
  return CRIMSON_OK;
}

int run_formation(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x,
                  double *d_res, int lhs, int mode, const AsmOut &o, double *d_qout, bool timed = true) {
  const unsigned T = 256;
  hipStream_t s = f->stream;
  const FormConst c = make_const(f->prm);
  if (lhs && f->npartial > 0 && !f->d_bc.p) {
    crimson::set_error("iBC holds 1- or 2-component velocity codes: BC is needed first (crimson_form_set_bc)");
    return CRIMSON_ESTATE;
  }
  if (f->prm.ideformwall && (!f->wall || !f->d_u)) {
    crimson::set_error(f->wall ? "ideformwall = 1: the displacement is needed first (crimson_form_set_displacement_dev)"
                               : "ideformwall = 1: the wall is needed first (crimson_form_set_wall)");
    return CRIMSON_ESTATE;
  }
  if (lhs) CR_TRY(ensure_order(f, d_x));
  if (timed) CR_HIP(hipEventRecord(f->ev[0], s));
  CR_TRY(crimson::form_flux_pass(f, d_y, d_ac, d_x, d_qout, nullptr));
  if (timed) CR_HIP(hipEventRecord(f->ev[1], s));
  // mixed precision: FP64 residual (e3 without tangent, res_node), FP32 tangent
  const bool t32 = lhs && f->prm.precision == 32;
  if (f->nel > 0) {
    if (lhs && !t32)
      e3_elem<true><<<crimson::grid_for(f->nel, kE3Threads), kE3Threads, 0, s>>>(
          f->nel, f->d_ien.p, f->d_node.p, c, f->ens(), f->d_rec.p, f->d_esc.p);
    else
      e3_elem<false><<<crimson::grid_for(f->nel, kE3Threads), kE3Threads, 0, s>>>(
          f->nel, f->d_ien.p, f->d_node.p, c, f->ens(), f->d_rec.p, f->d_esc.p);
    if (t32)
      e3_tangent32<<<crimson::grid_for(f->nel, T), T, 0, s>>>(f->nel, f->d_ien.p, f->d_node.p, c, f->ens(),
                                                              (float)f->prm.aspect_fp64, f->d_rec32.p);
  }
  CR_HIP(hipGetLastError());
  if (timed) CR_HIP(hipEventRecord(f->ev[2], s));
  const unsigned gn = crimson::grid_for((size_t)f->nshg * kNodeLanes, T);
  const unsigned gr = crimson::grid_for((size_t)f->nshg * kRowLanes, T);
  const int runs = f->sched == kSchedNumbering;  // slot i is row i: a wave's rows are consecutive
  if (t32) {
    if (mode == 0) {
      res_node<0><<<gn, T, 0, s>>>(f->nshg, f->d_inc_ptr.p, f->d_inc.p, f->d_rec.p, d_res);
      asm_row32<0><<<gr, T, 0, s>>>(f->nshg, f->d_slot.p, f->d_inc.p, f->d_lst_ptr.p, f->d_lst.p,
                                    f->d_rec32.p, o, f->asm_xcd);
    } else {
      res_node<1><<<gn, T, 0, s>>>(f->nshg, f->d_inc_ptr.p, f->d_inc.p, f->d_rec.p, d_res);
      asm_row32<1><<<gr, T, 0, s>>>(f->nshg, f->d_slot.p, f->d_inc.p, f->d_lst_ptr.p, f->d_lst.p,
                                    f->d_rec32.p, o, f->asm_xcd);
    }
  } else if (!lhs) {
    if (mode == 0)
      res_node<0><<<gn, T, 0, s>>>(f->nshg, f->d_inc_ptr.p, f->d_inc.p, f->d_rec.p, d_res);
    else
      res_node<1><<<gn, T, 0, s>>>(f->nshg, f->d_inc_ptr.p, f->d_inc.p, f->d_rec.p, d_res);
  } else if (mode == 0) {
    if (f->asm_chunk == 8)
      asm_row<0, 8><<<gr, T, 0, s>>>(f->nshg, f->d_slot.p, f->d_inc.p, f->d_ienp.p, f->d_xu.p, f->d_lst_ptr.p,
                                     f->d_lst.p, f->d_rec.p, f->d_esc.p, d_res, o, f->asm_xcd, runs);
    else
      asm_row<0><<<gr, T, 0, s>>>(f->nshg, f->d_slot.p, f->d_inc.p, f->d_ienp.p, f->d_xu.p, f->d_lst_ptr.p,
                                  f->d_lst.p, f->d_rec.p, f->d_esc.p, d_res, o, f->asm_xcd, runs);
  } else {
    if (f->asm_chunk == 8)
      asm_row<1, 8><<<gr, T, 0, s>>>(f->nshg, f->d_slot.p, f->d_inc.p, f->d_ienp.p, f->d_xu.p, f->d_lst_ptr.p,
                                     f->d_lst.p, f->d_rec.p, f->d_esc.p, d_res, o, f->asm_xcd, runs);
    else
      asm_row<1><<<gr, T, 0, s>>>(f->nshg, f->d_slot.p, f->d_inc.p, f->d_ienp.p, f->d_xu.p, f->d_lst_ptr.p,
                                  f->d_lst.p, f->d_rec.p, f->d_esc.p, d_res, o, f->asm_xcd, runs);
  }
  // the wall's rows and tangent (AsBMFG's xKebe goes through bc3lhs too)
  CR_TRY(crimson::form_wall_pass(f, lhs, mode, o.lhsK, o.V, mode == 1 && t32 ? o.V32 : nullptr, o.nnz));
  if (lhs && f->nbcl > 0) {
    {
      const unsigned gb = crimson::grid_for(f->nbcl, 256);
      if (mode == 0)
        bc3lhs_kernel<0, double><<<gb, 256, 0, s>>>(f->nbcl, f->d_bcl.p, f->d_bcn.p, o.lhsK, o.nnz);
      else if (t32)
        bc3lhs_kernel<1, float><<<gb, 256, 0, s>>>(f->nbcl, f->d_bcl.p, f->d_bcn.p, o.V32, o.nnz);
      else
        bc3lhs_kernel<1, double><<<gb, 256, 0, s>>>(f->nbcl, f->d_bcl.p, f->d_bcn.p, o.V, o.nnz);
    }
  }
  CR_HIP(hipGetLastError());
  if (timed) CR_HIP(hipEventRecord(f->ev[3], s));
  CR_TRY(crimson::form_boundary_pass(f, d_res, mode));  // the boundary loop after the interior (elmgmr.f90:325)
  if (timed) CR_HIP(hipEventRecord(f->ev[4], s));
  if (!timed) return CRIMSON_OK;  // enqueued only (graph capture)
  CR_HIP(hipStreamSynchronize(s));
  CR_HIP(hipEventElapsedTime(&f->ms[0], f->ev[0], f->ev[1]));
  CR_HIP(hipEventElapsedTime(&f->ms[1], f->ev[1], f->ev[2]));
  CR_HIP(hipEventElapsedTime(&f->ms[2], f->ev[2], f->ev[3]));
  CR_HIP(hipEventElapsedTime(&f->ms[3], f->ev[0], f->ev[4]));
  return CRIMSON_OK;
}

}  // namespace

namespace crimson {
// The flux pass (node_pack + flux_node) on the handle's stream: fills the
// packed nodes and the node records; q(nshg,9) and 1/rmass(nshg) when
// requested.
int form_flux_pass(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x, double *d_q,
                   double *d_invm) {
  const unsigned T = 256;
  hipStream_t s = f->stream;
  node_pack<<<crimson::grid_for(f->nshg, T), T, 0, s>>>(f->nshg, d_x, f->numnp, d_y, f->d_xu.p);
  flux_node<<<crimson::grid_for((size_t)f->nshg * kNodeLanes, T), T, 0, s>>>(
      f->nshg, f->d_inc_ptr.p, f->d_inc.p, f->d_ienp.p, f->d_xu.p, f->prm.mu, f->ens(), d_x, f->numnp, d_y, d_ac,
      f->d_node.p, d_q, d_invm);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}
hipStream_t form_stream(const crimson_form *f) { return f->stream; }
crimson_form_params form_params_of(const crimson_form *f) { return f->prm; }
int form_nshg(const crimson_form *f) { return f->nshg; }
}  // namespace crimson

extern "C" int crimson_form_setup(int nshg, int numnp, int nel, const int *ien, const int *colm,
                                  const int *rowp, int nnz, const int *iBC, const crimson_form_params *prm,
                                  crimson_form **out) {
  if (!out || !ien || !colm || !rowp || nshg <= 0 || numnp < nshg || nel < 0 || nnz <= 0) {
    crimson::set_error("crimson_form_setup: invalid sizes or null pointers (nshg=%d numnp=%d nel=%d nnz=%d)",
                       nshg, numnp, nel, nnz);
    return CRIMSON_EARG;
  }
  if ((long long)nel * 16 >= (1LL << 31)) {
    crimson::set_error("crimson_form_setup: nel=%d exceeds the 2^27 element limit of the packed maps", nel);
    return CRIMSON_EARG;
  }
  CR_TRY(validate_params(prm));
  // CSR (0-based) and checks
  std::vector<int> rp(nshg + 1), cols(nnz);
  for (int i = 0; i <= nshg; ++i) rp[i] = colm[i] - 1;
  if (rp[0] != 0 || rp[nshg] != nnz) {
    crimson::set_error("colm must start at 1 and end at nnz+1 (got %d .. %d, nnz=%d)", colm[0], colm[nshg], nnz);
    return CRIMSON_EARG;
  }
  for (int i = 0; i < nshg; ++i) {
    if (rp[i + 1] < rp[i]) {
      crimson::set_error("colm not monotone at row %d", i + 1);
      return CRIMSON_EARG;
    }
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      cols[k] = rowp[k] - 1;
      if (cols[k] < 0 || cols[k] >= nshg || (k > rp[i] && cols[k] <= cols[k - 1])) {
        crimson::set_error("rowp row %d not strictly ascending within [1,nshg]", i + 1);
        return CRIMSON_EARG;
      }
    }
  }
  std::vector<int> diagk(nshg);
  for (int i = 0; i < nshg; ++i) {
    diagk[i] = find_col(rp, cols, i, i);
    if (diagk[i] < 0) {
      crimson::set_error("row %d has no diagonal block", i + 1);
      return CRIMSON_ENODIAG;
    }
  }
  int npartial = 0;
  if (iBC) {
    for (int i = 0; i < nshg; ++i) {
      const int vb = (iBC[i] >> 3) & 7;
      npartial += vb != 0 && vb != 7;
      if ((iBC[i] >> 10) & 1) {
        crimson::set_error("node %d: periodic BC (iBC bit 10) is not supported", i + 1);
        return CRIMSON_EUNSUPPORTED;
      }
    }
    // The reference's code-6 row operations (bc3lhs.f90:648-680) do not
    // commute with another node's column operations, so their sum over the
    // elements depends on each element's local node order: not supported
    // where a code-6 node shares an element with another partial node.
    for (int e = 0; e < nel && npartial; ++e) {
      int n6 = 0, np = 0;
      for (int a = 0; a < 4; ++a) {
        const int g = ien[(size_t)a * nel + e] - 1;
        if (g < 0 || g >= nshg) continue;  // reported below
        const int vb = (iBC[g] >> 3) & 7;
        n6 += vb == 6;
        np += vb != 0 && vb != 7;
      }
      if (n6 > 0 && np > 1) {
        crimson::set_error("element %d: a node with velocity code 6 shares the element with another 1- or "
                           "2-component node; the reference's code-6 projection then depends on the local "
                           "node order (not supported)",
                           e + 1);
        return CRIMSON_EUNSUPPORTED;
      }
    }
  }
  // Row schedule of the assembly: the caller's numbering when it is already
  // local (99% of the elements span at most kLocalSpan node numbers, so that a
  // record line's second use stays within the last-level cache), else Morton
  // order of the coordinates (ensure_order).
  int far = 0;
  for (int e = 0; e < nel; ++e) {
    int lo = ien[e], hi = ien[e];
    for (int a = 1; a < 4; ++a) {
      lo = std::min(lo, ien[(size_t)a * nel + e]);
      hi = std::max(hi, ien[(size_t)a * nel + e]);
    }
    far += hi - lo > kLocalSpan;
  }
  const bool natural_sched = (long long)far * 100 <= (long long)nel;
  // node incidences in element order
  std::vector<int> inc_ptr(nshg + 1, 0), inc((size_t)nel * 4);
  for (int a = 0; a < 4; ++a)
    for (int e = 0; e < nel; ++e) {
      const int n = ien[(size_t)a * nel + e] - 1;
      if (n < 0 || n >= nshg) {
        crimson::set_error("ien(%d,%d)=%d out of range [1,%d]", e + 1, a + 1, n + 1, nshg);
        return CRIMSON_EARG;
      }
      inc_ptr[n + 1]++;
    }
  for (int i = 0; i < nshg; ++i) inc_ptr[i + 1] += inc_ptr[i];
  {
    std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int e = 0; e < nel; ++e)
      for (int a = 0; a < 4; ++a) {
        const int n = ien[(size_t)a * nel + e] - 1;
        inc[fill[n]++] = e * 4 + a;
      }
  }
  // Contribution lists of the off-diagonal blocks: for block k of row i, the
  // entries (t << 4 | a << 2 | b) of the row's incidences t = (e,a) whose node
  // ien(e,b) is the block's column, in element order (sparseloc,
  // fillsparse.f90:41, resolved once).
  std::vector<int> lst_ptr(nnz + 1, 0);
  for (int i = 0; i < nshg; ++i) {
    if (inc_ptr[i + 1] - inc_ptr[i] >= (1 << 12)) {
      crimson::set_error("node %d belongs to %d elements; at most 4095 supported", i + 1, inc_ptr[i + 1] - inc_ptr[i]);
      return CRIMSON_EUNSUPPORTED;
    }
    for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) {
      const int e = inc[p] >> 2, a = inc[p] & 3;
      for (int b = 0; b < 4; ++b) {
        if (b == a) continue;
        const int j = ien[(size_t)b * nel + e] - 1;
        const int k = find_col(rp, cols, i, j);
        if (k < 0) {
          crimson::set_error("element %d: pair (%d,%d) missing from the CSR graph", e + 1, i + 1, j + 1);
          return CRIMSON_EGRAPH;
        }
        lst_ptr[k + 1]++;
      }
    }
  }
  for (int k = 0; k < nnz; ++k) lst_ptr[k + 1] += lst_ptr[k];
  // two spare entries: the assembly reads the list in whole dwords
  std::vector<unsigned short> lst((size_t)lst_ptr[nnz] + 2, 0);
  {
    std::vector<int> fill(lst_ptr.begin(), lst_ptr.end() - 1);
    for (int i = 0; i < nshg; ++i)
      for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) {
        const int e = inc[p] >> 2, a = inc[p] & 3;
        for (int b = 0; b < 4; ++b) {
          if (b == a) continue;
          const int k = find_col(rp, cols, i, ien[(size_t)b * nel + e] - 1);
          lst[fill[k]++] = (unsigned short)(((p - inc_ptr[i]) << 4) | (a << 2) | b);
        }
      }
  }
  for (int e = 0; e < nel; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = a + 1; b < 4; ++b)
        if (ien[(size_t)a * nel + e] == ien[(size_t)b * nel + e]) {
          crimson::set_error("element %d repeats node %d", e + 1, ien[(size_t)a * nel + e]);
          return CRIMSON_EARG;
        }

  crimson_form *f = new crimson_form();
  f->sched = natural_sched ? kSchedNumbering : kSchedMorton;
  f->nshg = nshg;
  f->numnp = numnp;
  f->nel = nel;
  f->nnz = nnz;
  f->prm = *prm;
  int rc = CRIMSON_OK;
  auto fail = [&](int code) {
    crimson_form_free(f);
    return code;
  };
  if (hipGetDevice(&f->device) != hipSuccess) f->device = 0;
  if (hipStreamCreate(&f->stream) != hipSuccess) {
    crimson::set_error("hipStreamCreate failed (is a GPU visible?)");
    f->stream = nullptr;
    return fail(CRIMSON_EHIP);
  }
  for (auto &ev : f->ev)
    if (hipEventCreate(&ev) != hipSuccess) {
      crimson::set_error("hipEventCreate failed");
      return fail(CRIMSON_EHIP);
    }
  std::vector<int> ien0((size_t)nel * 4);
  for (size_t k = 0; k < ien0.size(); ++k) ien0[k] = ien[k] - 1;
  // per incidence, its element's nodes (the gathers of the flux pass and of
  // the assembly's shape gradients, read per row as one contiguous run).  One
  // spare entry (node 0) when there are no elements, so that every read of
  // entry 0 is in bounds.
  std::vector<int4> ienp(std::max<size_t>(inc.size(), 1), make_int4(0, 0, 0, 0));
  for (size_t p = 0; p < inc.size(); ++p) {
    const int e = inc[p] >> 2;
    ienp[p] = make_int4(ien0[e], ien0[(size_t)nel + e], ien0[2 * (size_t)nel + e], ien0[3 * (size_t)nel + e]);
  }
  if (inc.empty()) inc.push_back(0);  // the same spare entry (element 0's records are always allocated)
  hipStream_t s = f->stream;
  if ((rc = f->d_ien.upload(ien0.data(), ien0.size(), s)) ||
      (rc = f->d_inc_ptr.upload(inc_ptr.data(), inc_ptr.size(), s)) ||
      (rc = f->d_inc.upload(inc.data(), inc.size(), s)) ||
      (rc = f->d_diagk.upload(diagk.data(), diagk.size(), s)) ||
      (rc = f->d_rowptr.upload(rp.data(), rp.size(), s)) ||
      (rc = f->d_lst_ptr.upload(lst_ptr.data(), lst_ptr.size(), s)) ||
      (rc = f->d_lst.upload(lst.data(), lst.size(), s)) ||
      (rc = f->d_ienp.upload(ienp.data(), ienp.size(), s)) || (rc = f->d_xu.alloc((size_t)nshg * kXU)) ||
      (rc = f->d_node.alloc((size_t)nshg * kNRec)) || (rc = f->d_q.alloc((size_t)nshg * 9)) ||
      (rc = f->d_rec.alloc((size_t)std::max(nel, 1) * kRec)) ||
      (rc = f->d_esc.alloc((size_t)std::max(nel, 1) * kESc)))
    return fail(rc);
  if (prm->precision == 32 && (rc = f->d_rec32.alloc((size_t)std::max(nel, 1) * kRec32))) return fail(rc);
  if (iBC) {
    f->has_ibc = true;
    if ((rc = f->d_ibc.upload(iBC, nshg, s))) return fail(rc);
  }
  f->h_wrp = rp;
  f->h_wcols = cols;
  f->npartial = npartial;
  if (npartial > 0) {
    f->h_ibc.assign(iBC, iBC + nshg);
    f->h_rp = rp;
    f->h_cols = cols;
    f->h_ninc.resize(nshg);
    for (int i = 0; i < nshg; ++i) f->h_ninc[i] = inc_ptr[i + 1] - inc_ptr[i];
    f->h_nbnd.assign(nshg, 0);
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    crimson::set_error("upload of the formation maps failed");
    return fail(CRIMSON_EHIP);
  }
  *out = f;
  return CRIMSON_OK;
}

extern "C" int crimson_form_free(crimson_form *f) {
  if (!f) return CRIMSON_OK;
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  for (auto &ev : f->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
  return CRIMSON_OK;
}

extern "C" int crimson_form_interior_dev(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x,
                                         double *d_res, double *d_lhsK, double *d_lhsP, double *d_qres, int lhs) {
  if (!f || !d_y || !d_ac || !d_x || !d_res || (lhs && (!d_lhsK || !d_lhsP))) {
    crimson::set_error("crimson_form_interior_dev: null argument");
    return CRIMSON_EARG;
  }
  AsmOut o{d_lhsK, d_lhsP, nullptr, nullptr, (size_t)f->nnz};
  return run_formation(f, d_y, d_ac, d_x, d_res, lhs ? 1 : 0, 0, o, d_qres);
}

extern "C" int crimson_form_interior(crimson_form *f, const double *y, const double *ac, const double *x,
                                     double *res, double *lhsK, double *lhsP, double *qres, int lhs) {
  if (!f || !y || !ac || !x || !res || (lhs && (!lhsK || !lhsP))) {
    crimson::set_error("crimson_form_interior: null argument");
    return CRIMSON_EARG;
  }
  hipStream_t s = f->stream;
  const size_t N = f->nshg;
  CR_TRY(f->s_y.upload(y, N * 4, s));
  CR_TRY(f->s_ac.upload(ac, N * 4, s));
  CR_TRY(f->s_x.upload(x, (size_t)f->numnp * 3, s));
  CR_TRY(f->s_res.alloc(N * 4));
  if (lhs) {
    CR_TRY(f->s_K.alloc((size_t)f->nnz * 9));
    CR_TRY(f->s_P.alloc((size_t)f->nnz * 4));
  }
  AsmOut o{f->s_K.p, f->s_P.p, nullptr, nullptr, (size_t)f->nnz};
  CR_TRY(run_formation(f, f->s_y.p, f->s_ac.p, f->s_x.p, f->s_res.p, lhs ? 1 : 0, 0, o, qres ? f->d_q.p : nullptr));
  CR_HIP(hipMemcpyAsync(res, f->s_res.p, N * 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  if (lhs) {
    CR_HIP(hipMemcpyAsync(lhsK, f->s_K.p, (size_t)f->nnz * 9 * sizeof(double), hipMemcpyDeviceToHost, s));
    CR_HIP(hipMemcpyAsync(lhsP, f->s_P.p, (size_t)f->nnz * 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (qres) CR_HIP(hipMemcpyAsync(qres, f->d_q.p, N * 9 * sizeof(double), hipMemcpyDeviceToHost, s));
  CR_HIP(hipStreamSynchronize(s));
  return CRIMSON_OK;
}

extern "C" int crimson_form_interior_to_lhs(crimson_form *f, const double *d_y, const double *d_ac,
                                            const double *d_x, double *d_res, crimson_lhs *lhs) {
  if (!f || !lhs || !d_y || !d_ac || !d_x || !d_res) {
    crimson::set_error("crimson_form_interior_to_lhs: null argument");
    return CRIMSON_EARG;
  }
  if (lhs->nNo != f->nshg || lhs->nnz != f->nnz) {
    crimson::set_error("lhs (nNo=%d nnz=%d) does not match the formation graph (nshg=%d nnz=%d)", lhs->nNo,
                       lhs->nnz, f->nshg, f->nnz);
    return CRIMSON_EARG;
  }
  lhs->vdof = 4;
  lhs->gt_is_d = true;  // Gt = -mG^T = mD for the solfar packing
  if (f->prm.precision == 32) {
    CR_TRY(lhs->d_V32.alloc((size_t)16 * f->nnz));
    lhs->prec = 32;
    AsmOut o{nullptr, nullptr, nullptr, lhs->d_V32.p, (size_t)f->nnz};
    return run_formation(f, d_y, d_ac, d_x, d_res, 1, 1, o, nullptr);
  }
  CR_TRY(lhs->d_V.alloc((size_t)16 * f->nnz));
  lhs->prec = 64;
  AsmOut o{nullptr, nullptr, lhs->d_V.p, nullptr, (size_t)f->nnz};
  return run_formation(f, d_y, d_ac, d_x, d_res, 1, 1, o, nullptr);
}

// Launch configurations of the formation chosen by timing them on the
// caller's arrays: the assembly's row schedule (the caller's numbering,
// Morton order of the nodes, or runs of consecutive rows in Morton order of
// their centroids), its workgroups' dealing to the XCDs (a contiguous range
// each, or turns of kAsmXcdChunk) and, for the FP64 assembly, its incidence
// chunk (16 or 8).  No candidate changes a result: schedule and dealing
// order rows, never sums, and both chunk sizes give incidence t to lane
// t % 16.
constexpr unsigned kAsmXcdChunk = 64;
extern "C" int crimson_form_autotune(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x,
                                     double *d_res, crimson_lhs *lhs, int reps, int mask, int *choice,
                                     double *ms_out) {
  if (!f || !lhs || !d_y || !d_ac || !d_x || !d_res || reps < 1) {
    crimson::set_error("crimson_form_autotune: invalid arguments");
    return CRIMSON_EARG;
  }
  const bool fp32 = f->prm.precision == 32;
  // candidate c: schedule c % 3, 8-incidence chunks (c / 3) % 2, XCD turns c / 6;
  // the FP32 assembly has no chunk choice
  const int ncand = 12, valid = fp32 ? 0x1c7 : 0xfff;
  if (mask == 0) mask = valid;
  if (mask & ~valid) {
    crimson::set_error("crimson_form_autotune: candidate mask names a configuration this handle lacks");
    return CRIMSON_EARG;
  }
  double best = 1e300;
  int bestc = -1;
  for (int cand = 0; cand < ncand; ++cand) {
    if (!(mask & (1 << cand))) continue;
    const int sched = cand % 3;
    f->asm_chunk = (cand / 3) % 2 ? 8 : 16;
    f->asm_xcd = cand / 6 ? kAsmXcdChunk : 0;
    if (f->sched != sched || !f->d_order.p) {
      f->sched = sched;
      f->d_order.release();
      CR_TRY(ensure_order(f, d_x));
    }
    double t = 0.0;
    for (int r = 0; r <= reps; ++r) {  // the first call is a warm-up
      CR_TRY(crimson_form_interior_to_lhs(f, d_y, d_ac, d_x, d_res, lhs));
      if (r > 0) t += f->ms[2];
    }
    t /= reps;
    if (ms_out) ms_out[cand] = t;
    if (bestc < 0 || t < best) {
      best = t;
      bestc = cand;
    }
  }
  f->asm_chunk = (bestc / 3) % 2 ? 8 : 16;
  f->asm_xcd = bestc / 6 ? kAsmXcdChunk : 0;
  if (f->sched != bestc % 3) {
    f->sched = bestc % 3;
    f->d_order.release();
    CR_TRY(ensure_order(f, d_x));
  }
  if (choice) *choice = bestc;
  return CRIMSON_OK;
}

extern "C" int crimson_form_set_instances(crimson_form *f, const unsigned short *inst, int ninst, const double *mu) {
  if (!f || ninst < 0 || (ninst > 0 && (!inst || !mu))) {
    crimson::set_error("crimson_form_set_instances: invalid arguments");
    return CRIMSON_EARG;
  }
  if (ninst == 0) {
    f->ninst = 0;
    return CRIMSON_OK;
  }
  for (int k = 0; k < ninst; ++k)
    if (!(mu[k] >= 0.0)) {
      crimson::set_error("instance %d: viscosity %g", k, mu[k]);
      return CRIMSON_EARG;
    }
  for (int e = 0; e < f->nel; ++e)
    if (inst[e] >= ninst) {
      crimson::set_error("element %d: instance %d of %d", e + 1, (int)inst[e], ninst);
      return CRIMSON_EARG;
    }
  CR_TRY(f->d_inst.upload(inst, (size_t)std::max(f->nel, 1), f->stream));
  CR_TRY(f->d_mu_inst.upload(mu, (size_t)ninst, f->stream));
  CR_HIP(hipStreamSynchronize(f->stream));
  f->ninst = ninst;
  return crimson::form_boundary_instances(f);
}

namespace crimson {
// What a captured formation graph baked in: the allocation serial of every
// buffer its kernels read or write, and the scalars that chose its kernels
// and launch shapes.  A replay is refused once any of them has changed.
std::vector<unsigned long long> graph_signature(const crimson_form *f, const crimson_lhs *lhs) {
  std::vector<unsigned long long> v = {
      f->d_ien.serial, f->d_inc_ptr.serial, f->d_inc.serial, f->d_diagk.serial, f->d_rowptr.serial,
      f->d_ibc.serial, f->d_lst_ptr.serial, f->d_order.serial, f->d_slot.serial, f->d_lst.serial,
      f->d_ienp.serial, f->d_xu.serial, f->d_node.serial, f->d_rec.serial, f->d_esc.serial,
      f->d_rec32.serial, f->d_bc.serial,
      f->d_bcl.serial, f->d_bcn.serial, f->d_inst.serial, f->d_mu_inst.serial, f->d_binst.serial,
      f->d_ienb.serial, f->d_bcode.serial, f->d_bcb.serial, f->d_permit.serial, f->d_brec.serial,
      f->d_bflx.serial, f->d_sptr.serial, f->d_selem.serial, f->d_bnode.serial, f->d_binc_ptr.serial,
      f->d_binc.serial, lhs->d_V.serial, lhs->d_V32.serial, f->wall_gen, (unsigned long long)f->wall,
      (unsigned long long)(uintptr_t)f->d_u, f->d_wface.serial, f->d_wtag.serial, f->d_uref.serial,
      f->d_wres.serial, f->d_wrec.serial, f->d_wk.serial, f->d_wptr.serial, f->d_wsrc.serial,
      (unsigned long long)f->ninst, (unsigned long long)f->nelb, (unsigned long long)f->npartial,
      (unsigned long long)f->nbcl, (unsigned long long)f->prm.precision, (unsigned long long)f->asm_chunk,
      (unsigned long long)f->asm_xcd, (unsigned long long)f->sched, (unsigned long long)lhs->vdof,
      (unsigned long long)lhs->prec, (unsigned long long)lhs->gt_is_d};
  return v;
}

int form_prepare_to_lhs(crimson_form *f, const double *d_x, crimson_lhs *lhs) {
  if (lhs->nNo != f->nshg || lhs->nnz != f->nnz) {
    crimson::set_error("lhs (nNo=%d nnz=%d) does not match the formation graph (nshg=%d nnz=%d)", lhs->nNo,
                       lhs->nnz, f->nshg, f->nnz);
    return CRIMSON_EARG;
  }
  if (f->npartial > 0 && !f->d_bc.p) {
    crimson::set_error("iBC holds 1- or 2-component velocity codes: BC is needed first (crimson_form_set_bc)");
    return CRIMSON_ESTATE;
  }
  lhs->vdof = 4;
  lhs->gt_is_d = true;
  if (f->prm.precision == 32) {
    CR_TRY(lhs->d_V32.alloc((size_t)16 * f->nnz));
    lhs->prec = 32;
  } else {
    CR_TRY(lhs->d_V.alloc((size_t)16 * f->nnz));
    lhs->prec = 64;
  }
  return ensure_order(f, d_x);
}

int form_enqueue_to_lhs(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x, double *d_Res4,
                        crimson_lhs *lhs) {
  AsmOut o{nullptr, nullptr, nullptr, nullptr, (size_t)f->nnz};
  if (f->prm.precision == 32)
    o.V32 = lhs->d_V32.p;
  else
    o.V = lhs->d_V.p;
  return run_formation(f, d_y, d_ac, d_x, d_Res4, 1, 1, o, nullptr, false);
}
}  // namespace crimson

extern "C" int crimson_form_graph_create(crimson_form *f, const double *d_y, const double *d_ac, const double *d_x,
                                         double *d_res, crimson_lhs *lhs, crimson_graph **out) {
  if (!f || !lhs || !d_y || !d_ac || !d_x || !d_res || !out) {
    crimson::set_error("crimson_form_graph_create: null argument");
    return CRIMSON_EARG;
  }
  if (lhs->nNo != f->nshg || lhs->nnz != f->nnz) {
    crimson::set_error("lhs (nNo=%d nnz=%d) does not match the formation graph (nshg=%d nnz=%d)", lhs->nNo,
                       lhs->nnz, f->nshg, f->nnz);
    return CRIMSON_EARG;
  }
  // everything that allocates or synchronizes happens before the capture
  lhs->vdof = 4;
  lhs->gt_is_d = true;
  AsmOut o{nullptr, nullptr, nullptr, nullptr, (size_t)f->nnz};
  if (f->prm.precision == 32) {
    CR_TRY(lhs->d_V32.alloc((size_t)16 * f->nnz));
    lhs->prec = 32;
    o.V32 = lhs->d_V32.p;
  } else {
    CR_TRY(lhs->d_V.alloc((size_t)16 * f->nnz));
    lhs->prec = 64;
    o.V = lhs->d_V.p;
  }
  CR_TRY(ensure_order(f, d_x));
  crimson_graph *gr = new crimson_graph();
  gr->s = f->stream;
  if (hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    delete gr;
    crimson::set_error("hipStreamBeginCapture failed");
    return CRIMSON_EHIP;
  }
  const int rc = run_formation(f, d_y, d_ac, d_x, d_res, 1, 1, o, nullptr, false);
  const hipError_t ec = hipStreamEndCapture(f->stream, &gr->g);
  if (rc || ec != hipSuccess || hipGraphInstantiate(&gr->exec, gr->g, nullptr, nullptr, 0) != hipSuccess) {
    if (gr->g) (void)hipGraphDestroy(gr->g);
    delete gr;
    if (!rc) crimson::set_error("capture of the formation into a HIP graph failed");
    return rc ? rc : CRIMSON_EHIP;
  }
  gr->f = f;
  gr->lhs = lhs;
  gr->sig = crimson::graph_signature(f, lhs);
  *out = gr;
  return CRIMSON_OK;
}

extern "C" int crimson_graph_launch(crimson_graph *g) {
  if (!g) {
    crimson::set_error("crimson_graph_launch: null graph");
    return CRIMSON_EARG;
  }
  std::vector<unsigned long long> sig;
  if (g->f) sig = crimson::graph_signature(g->f, g->lhs);
  if (g->extra) {
    const std::vector<unsigned long long> e = g->extra->signature();
    sig.insert(sig.end(), e.begin(), e.end());
  }
  if (sig != g->sig) {
    crimson::set_error("crimson_graph_launch: a buffer or setting the graph captured has changed since "
                       "crimson_form_graph_create (instances, boundary, BC, precision, schedule or the matrix); "
                       "free the graph and capture it again");
    return CRIMSON_ESTATE;
  }
  if (g->segx.empty()) {
    CR_HIP(hipGraphLaunch(g->exec, g->s));
  } else {
    for (hipGraphExec_t x : g->segx) CR_HIP(hipGraphLaunch(x, g->s));
  }
  CR_HIP(hipStreamSynchronize(g->s));
  if (g->extra) return g->extra->after_launch();
  return CRIMSON_OK;
}

extern "C" int crimson_graph_free(crimson_graph *g) {
  if (!g) return CRIMSON_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->g) (void)hipGraphDestroy(g->g);
  for (hipGraphExec_t x : g->segx) (void)hipGraphExecDestroy(x);
  for (hipGraph_t h : g->seg) (void)hipGraphDestroy(h);
  delete g->extra;
  delete g;
  return CRIMSON_OK;
}

namespace crimson {
int form_bc3lhs_table(crimson_form *f) {
  if (f->npartial == 0 || !f->d_bc.p) return CRIMSON_OK;
  const int n = f->nshg;
  std::vector<int> pidx(n, -1);
  std::vector<BclNode> nodes;
  const std::vector<double> &hbc = f->h_bc;
  for (int i = 0; i < n; ++i) {
    const int vb = (f->h_ibc[i] >> 3) & 7;
    if (vb == 0 || vb == 7) continue;
    pidx[i] = (int)nodes.size();
    BclNode nd;
    nd.code = vb;
    nd.count = f->h_ninc[i] + f->h_nbnd[i];
    for (int k = 0; k < 3; ++k) nd.b[k] = hbc[i + (size_t)n * (3 + k)];
    nodes.push_back(nd);
  }
  std::vector<int4> blk;
  for (int i = 0; i < n; ++i)
    for (int k = f->h_rp[i]; k < f->h_rp[i + 1]; ++k) {
      const int j = f->h_cols[k];
      if (pidx[i] < 0 && pidx[j] < 0) continue;
      blk.push_back(make_int4(k, pidx[i], pidx[j], i == j ? 1 : 0));
    }
  CR_TRY(f->d_bcn.upload(nodes.data(), nodes.size(), f->stream));
  CR_TRY(f->d_bcl.upload(blk.data(), blk.size(), f->stream));
  CR_HIP(hipStreamSynchronize(f->stream));
  f->nbcl = (int)blk.size();
  return CRIMSON_OK;
}
}  // namespace crimson

extern "C" int crimson_form_set_bc(crimson_form *f, const double *BC, int ndofBC) {
  if (!f || (!BC && ndofBC > 0) || ndofBC < 0) {
    crimson::set_error("crimson_form_set_bc: invalid arguments");
    return CRIMSON_EARG;
  }
  if (ndofBC == 0) {
    f->d_bc.release();
    f->d_bcl.release();
    f->d_bcn.release();
    f->nbcl = 0;
    f->ndofbc = 0;
    f->h_bc.clear();
    return CRIMSON_OK;
  }
  if (f->npartial > 0 && ndofBC < 6) {
    crimson::set_error("crimson_form_set_bc: %d nodes have 1- or 2-component velocity codes, which read "
                       "BC(:,4:6); ndofBC = %d",
                       f->npartial, ndofBC);
    return CRIMSON_EARG;
  }
  CR_TRY(f->d_bc.upload(BC, (size_t)f->nshg * ndofBC, f->stream));
  f->ndofbc = ndofBC;
  if (f->npartial > 0) f->h_bc.assign(BC, BC + (size_t)f->nshg * ndofBC);
  return crimson::form_bc3lhs_table(f);
}

extern "C" int crimson_form_bc3res_dev(crimson_form *f, double *d_res, int res4_layout) {
  if (!f || !d_res) {
    crimson::set_error("crimson_form_bc3res_dev: null argument");
    return CRIMSON_EARG;
  }
  if (!f->has_ibc) return CRIMSON_OK;
  if (f->npartial > 0 && !f->d_bc.p) {
    crimson::set_error("iBC holds 1- or 2-component velocity codes: BC is needed first (crimson_form_set_bc)");
    return CRIMSON_ESTATE;
  }
  CR_TRY(crimson::form_enqueue_bc3res(f, d_res, res4_layout));
  CR_HIP(hipStreamSynchronize(f->stream));
  return CRIMSON_OK;
}

namespace crimson {
int form_enqueue_bc3res(crimson_form *f, double *d_res, int res4_layout) {
  if (!f->has_ibc) return CRIMSON_OK;
  if (f->npartial > 0 && !f->d_bc.p) {
    crimson::set_error("iBC holds 1- or 2-component velocity codes: BC is needed first (crimson_form_set_bc)");
    return CRIMSON_ESTATE;
  }
  bc3res_kernel<<<crimson::grid_for(f->nshg, 256), 256, 0, f->stream>>>(f->nshg, f->d_ibc.p, f->d_bc.p, d_res,
                                                                         res4_layout ? 1 : 0);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}
}  // namespace crimson

extern "C" int crimson_form_timing(const crimson_form *f, double *out4) {
  if (!f || !out4) {
    crimson::set_error("crimson_form_timing: null argument");
    return CRIMSON_EARG;
  }
  for (int k = 0; k < 4; ++k) out4[k] = f->ms[k];
  return CRIMSON_OK;
}

extern "C" int crimson_genadj(int nshg, int nel, const int *ien, int *colm, int *rowp, int maxnnz, int *nnz_out) {
  if (nshg <= 0 || nel < 0 || !ien || !colm || !nnz_out) {
    crimson::set_error("crimson_genadj: invalid arguments");
    return CRIMSON_EARG;
  }
  // node -> element incidences, then per row the sorted unique union of the
  // incident elements' nodes (what Asadj + genadj's sort produce)
  std::vector<int> ptr(nshg + 1, 0);
  for (int a = 0; a < 4; ++a)
    for (int e = 0; e < nel; ++e) {
      const int n = ien[(size_t)a * nel + e];
      if (n < 1 || n > nshg) {
        crimson::set_error("ien(%d,%d)=%d out of range", e + 1, a + 1, n);
        return CRIMSON_EARG;
      }
      ptr[n]++;
    }
  for (int i = 0; i < nshg; ++i) ptr[i + 1] += ptr[i];
  std::vector<int> inc((size_t)ptr[nshg]);
  {
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < nel; ++e)
      for (int a = 0; a < 4; ++a) inc[fill[ien[(size_t)a * nel + e] - 1]++] = e;
  }
  std::vector<int> buf;
  long long nnz = 0;
  colm[0] = 1;
  for (int i = 0; i < nshg; ++i) {
    buf.clear();
    for (int p = ptr[i]; p < ptr[i + 1]; ++p)
      for (int b = 0; b < 4; ++b) buf.push_back(ien[(size_t)b * nel + inc[p]]);
    // a node no element references gets an empty row, as Asadj leaves it
    // (crimson_form_setup then rejects the row: it has no diagonal block)
    std::sort(buf.begin(), buf.end());
    const size_t len = std::unique(buf.begin(), buf.end()) - buf.begin();
    if (rowp && nnz + (long long)len <= maxnnz)
      std::copy(buf.begin(), buf.begin() + len, rowp + nnz);
    nnz += len;
    colm[i + 1] = (int)(nnz + 1);
  }
  *nnz_out = (int)nnz;
  if (nnz > maxnnz || !rowp) {
    crimson::set_error("crimson_genadj: need %lld column entries, capacity %d", nnz, maxnnz);
    return CRIMSON_EARG;
  }
  return CRIMSON_OK;
}
