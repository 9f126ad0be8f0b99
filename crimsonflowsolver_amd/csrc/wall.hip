// Deformable vessel walls (ideformwall = 1) of ElmGMR's boundary loop for P1
// tets: e3bvar's wall section (e3bvar.f90:462-1017) as e3b uses it
// (e3b.f90:374-395, 488-575), on the boundary faces with iBCB bit 4.
//
// Per wall face and boundary quadrature point the reference forms, in the
// face's flat frame (rows v1 = first edge, v2 = n x v1, v3 = n):
//   the membrane      rlKwall(a,i) = sum_m B(:,a,i) . D B_m (u_m + uref_m) WdetJb h
//   its tangent       LHSwall(3(i-1)+k, a, m) = B(:,a,i) . D B(:,m,k) WdetJb h
//                     iwallstiffactor betai Delt^2 alfi
//   the wall mass     vdot = rhovw h sum_a N_a ac_a
//   support, damping  usup1 = k sum_a N_a u_a, velsup = c sum_a N_a v_a
//   the ring tethering (first point only) along the end rings' edges
// and the diagonal tangent Wall_LHSfactor N_a N_b.  B is the in-plane strain
// (3 rows) and transverse shear (2 rows) of the P1 displacement; its in-plane
// gradients come from e3bvar's own construction, the inverse of the
// rotated face nodes' coordinates as a 3x3 "jacobian" (node 4 at the
// origin), which equals the barycentric gradients in the plane whenever the
// face's plane misses the origin (a plane through it divides by zero, in the
// reference as here).
//
//   wall_elem      one thread per wall face: the rows of the three face
//                  nodes (9 doubles) and, with the tangent, xKebe of the 9
//                  face-node pairs (81 doubles), summed over the points in
//                  the reference's order
//   wall_lhs_add   one thread per matrix block the wall touches: the sum of
//                  its face-pair contributions, in face order, added to the
//                  block before bc3lhs (lhsK(9,nnz) or the FP64/FP32 planes)
//   wall_rows_add  brec += the wall rows (after bnd_elem)
//   wall_predict / wall_correct / wall_alpha: the wall nodes' displacement
//                  in the generalized-alpha step (itrPC.f90)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"

#include "form_common.hpp"

namespace {

constexpr double kTriA = 0.3333333333333330000000000000000000;  // symtri.c:20
constexpr double kTriB = 0.666666666666667;                     // symtri.c:21
constexpr double kTriC = 0.166666666666667;                     // symtri.c:22

__device__ inline double tri_pt(int nqb, int q, int a) {
  if (nqb == 1) return kTriA;
  return q == a ? kTriB : kTriC;
}

template <bool LHS>
__global__ void __launch_bounds__(256) wall_elem(int nwf, const int *__restrict__ wface, int nelb,
                                                 const int *__restrict__ ienb, const double *__restrict__ node,
                                                 const double *__restrict__ du, const double *__restrict__ uref,
                                                 int nshg, const int *__restrict__ wtag, int nqb, WallConst wc,
                                                 double *__restrict__ wres, double *__restrict__ wrec) {
  const int w = xcd_block() * blockDim.x + threadIdx.x;
  if (w >= nwf) return;
  const int e = wface[w];
  double xl[3][3], vl[3][3], al[3][3], ul[3][3], us[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int g = ienb[a * (size_t)nelb + e];
    const double *r = node + (size_t)g * kNRec;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xl[a][i] = r[N_X + i];
      vl[a][i] = r[N_U + i];
      al[a][i] = r[N_A + i];
      ul[a][i] = du[g + (size_t)nshg * i];
      us[a][i] = ul[a][i] + (uref ? uref[g + (size_t)nshg * i] : 0.0);
    }
  }
  // the face normal and WdetJb (e3bvar.f90, lcsyst 1)
  double v1[3], v2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v1[i] = xl[1][i] - xl[0][i];
    v2[i] = xl[2][i] - xl[0][i];
  }
  double t1 = v1[1] * v2[2] - v2[1] * v1[2];
  double t2 = v2[0] * v1[2] - v1[0] * v2[2];
  double t3 = v1[0] * v2[1] - v2[0] * v1[1];
  const double tn = 1.0 / sqrt(t1 * t1 + t2 * t2 + t3 * t3);
  const double n3[3] = {t1 * tn, t2 * tn, t3 * tn};
  const double W = (nqb == 1 ? 2.0 * 1.0 : 2.0 * kTriA) / (4.0 * tn);  // the same at every point
  // the flat frame: rot rows v1 (first edge), v2 = n x v1, n
  double rot[3][3];
  {
    const double t = 1.0 / sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) rot[0][i] = v1[i] * t;
    t1 = n3[1] * rot[0][2] - rot[0][1] * n3[2];
    t2 = rot[0][0] * n3[2] - n3[0] * rot[0][2];
    t3 = n3[0] * rot[0][1] - rot[0][0] * n3[1];
    const double s = 1.0 / sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    rot[1][0] = t1 * s;
    rot[1][1] = t2 * s;
    rot[1][2] = t3 * s;
#pragma unroll
    for (int i = 0; i < 3; ++i) rot[2][i] = n3[i];
  }
  // J(i,j) = 0.5 * (rotated node j+1)_i (shglb = dN/dxi / 2, node 4 at the
  // origin), its cofactor inverse D, and the in-plane gradients
  // sg(i,n) = sum_j D(j,i) shglb(j,n) of the face nodes (i = 0, 1)
  double J[3][3];
#pragma unroll
  for (int n = 0; n < 3; ++n)
#pragma unroll
    for (int i = 0; i < 3; ++i) J[i][n] = 0.5 * (rot[i][0] * xl[n][0] + rot[i][1] * xl[n][1] + rot[i][2] * xl[n][2]);
  double D[3][2];  // the inverse's first two columns: the in-plane directions
  const double c00 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  const double c01 = J[2][1] * J[0][2] - J[0][1] * J[2][2];
  const double c02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  const double td = 1.0 / (c00 * J[0][0] + c01 * J[1][0] + c02 * J[2][0]);
  D[0][0] = c00 * td;
  D[0][1] = c01 * td;
  D[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * td;
  D[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * td;
  D[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * td;
  D[2][1] = (J[2][0] * J[0][1] - J[0][0] * J[2][1]) * td;
  double sg[2][3];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int n = 0; n < 3; ++n) sg[i][n] = 0.5 * D[n][i];  // shglb(j,n) = 0.5 delta_jn for the face nodes
  // B(j, n, i) and the material matrix (isotropic: D7 D8 / D8 D7 / D9 / D10 / D10)
  double B[5][3][3];
#pragma unroll
  for (int n = 0; n < 3; ++n)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      B[0][n][i] = sg[0][n] * rot[0][i];
      B[1][n][i] = sg[1][n] * rot[1][i];
      B[2][n][i] = sg[1][n] * rot[0][i] + sg[0][n] * rot[1][i];
      B[3][n][i] = sg[0][n] * rot[2][i];
      B[4][n][i] = sg[1][n] * rot[2][i];
    }
  // the strain of u + uref summed over the face nodes, and D times it
  double st[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int j = 0; j < 5; ++j) st[j] += B[j][m][0] * us[m][0] + B[j][m][1] * us[m][1] + B[j][m][2] * us[m][2];
  const double Ds[5] = {wc.d7 * st[0] + wc.d8 * st[1], wc.d8 * st[0] + wc.d7 * st[1], wc.d9 * st[2],
                        wc.d10 * st[3], wc.d10 * st[4]};
  double rk[3][3];  // rlKwall / (W h): B^T D strain
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      rk[a][i] = B[0][a][i] * Ds[0] + B[1][a][i] * Ds[1] + B[2][a][i] * Ds[2] + B[3][a][i] * Ds[3] +
                 B[4][a][i] * Ds[4];
  // the ring tethering (e3bvar.f90:876-975), added at the first point
  double rr[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  int r0 = -1, r1 = -1;
  double dvec[3] = {0.0, 0.0, 0.0}, hf = 0.0;
  if (wtag && (wc.iring_k || wc.iring_c)) {
    int nr = 0;
    double rm[3] = {0.0, 0.0, 0.0}, xm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (wtag[3 * (size_t)w + a]) {
        if (nr == 0) r0 = a; else if (nr == 1) r1 = a;
        ++nr;
#pragma unroll
        for (int i = 0; i < 3; ++i) rm[i] += xl[a][i];
      }
    if (nr > 1) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        xm[i] = (xl[0][i] + xl[1][i] + xl[2][i]) / 3.0;
        rm[i] = rm[i] / nr;
      }
      double d[3], e2[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        d[i] = rm[i] - xm[i];
        e2[i] = xl[r1][i] - rm[i];
      }
      double nn = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
#pragma unroll
      for (int i = 0; i < 3; ++i) e2[i] = e2[i] / nn;
      const double dp = d[0] * e2[0] + d[1] * e2[1] + d[2] * e2[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) d[i] = d[i] - e2[i] * dp;
      nn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        dvec[i] = d[i] / nn;
        e2[i] = xl[r1][i] - xl[r0][i];
      }
      const double len = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]) / 6.0;
      double f1[3] = {0.0, 0.0, 0.0}, f2[3] = {0.0, 0.0, 0.0};
      if (wc.iring_k)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          f1[i] += wc.ring_k * ul[r0][i];
          f2[i] += wc.ring_k * ul[r1][i];
        }
      if (wc.iring_c)  // both ends take the first ring node's velocity, as e3bvar.f90:937-940 writes it
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          f1[i] += wc.ring_c * vl[r0][i];
          f2[i] += wc.ring_c * vl[r0][i];
        }
      const double p1 = dvec[0] * f1[0] + dvec[1] * f1[1] + dvec[2] * f1[2];
      const double p2 = dvec[0] * f2[0] + dvec[1] * f2[1] + dvec[2] * f2[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        rr[r0][i] = len * (2.0 * (dvec[i] * p1) + dvec[i] * p2);
        rr[r1][i] = len * (dvec[i] * p1 + 2.0 * (dvec[i] * p2));
      }
      if (wc.iring_k) hf += len * wc.ring_k * wc.ring_klhs;
      if (wc.iring_c) hf += len * wc.ring_c * wc.ring_clhs;
    } else {
      r0 = r1 = -1;
    }
  }
  // the rows: per point, rl -= N_a (vdot + velsup + usup1) W, then
  // rl -= rlKwall (the ring part at the first point only)
  double rl[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) rl[a][i] = 0.0;
  const double Wh = W * wc.h;
  for (int q = 0; q < nqb; ++q) {
    const double N[3] = {tri_pt(nqb, q, 0), tri_pt(nqb, q, 1), tri_pt(nqb, q, 2)};
    double rn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double acc = (N[0] * al[0][i] + N[1] * al[1][i] + N[2] * al[2][i]) * wc.h * wc.rhovw;
      if (wc.idamp) acc += (N[0] * vl[0][i] + N[1] * vl[1][i] + N[2] * vl[2][i]) * wc.damp;
      if (wc.isupp) acc += (N[0] * ul[0][i] + N[1] * ul[1][i] + N[2] * ul[2][i]) * wc.supp;
      rn[i] = acc * W;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        rl[a][i] = rl[a][i] - N[a] * rn[i];
        rl[a][i] = rl[a][i] - (rk[a][i] * Wh + (q == 0 ? rr[a][i] : 0.0));
      }
  }
  double *o = wres + (size_t)w * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) o[3 * a + i] = rl[a][i];
  if (!LHS) return;
  // xKebe(3(i-1)+k, a, b): per point the diagonal mass/support/damping term
  // then LHSwall (the same at every point but the ring, at the first)
  double wlf = wc.mass;
  if (wc.isupp) wlf += wc.supp_lhs;
  if (wc.idamp) wlf += wc.damp_lhs;
  wlf = wc.massf * wlf * W;
  const double kf = Wh * wc.kstiff;
  double *out = wrec + (size_t)w * 81;
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double K[9];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        // (D B)(:, b, k) for the three k
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double db0 = wc.d7 * B[0][b][k] + wc.d8 * B[1][b][k];
          const double db1 = wc.d8 * B[0][b][k] + wc.d7 * B[1][b][k];
          const double s = B[0][a][i] * db0 + B[1][a][i] * db1 + B[2][a][i] * (wc.d9 * B[2][b][k]) +
                           B[3][a][i] * (wc.d10 * B[3][b][k]) + B[4][a][i] * (wc.d10 * B[4][b][k]);
          double lw = s * kf;
          K[3 * i + k] = 0.0;
          double ring = 0.0;
          if (r0 >= 0) {
            const double c = dvec[i] * dvec[k] * hf;
            if (a == r0 && b == r0) ring = c * 2.0;
            else if ((a == r0 && b == r1) || (a == r1 && b == r0)) ring = c;
            else if (a == r1 && b == r1) ring = c * 2.0;
          }
          double acc = 0.0;
          for (int q = 0; q < nqb; ++q) {
            if (i == k) acc += wlf * tri_pt(nqb, q, a) * tri_pt(nqb, q, b);
            acc += lw + (q == 0 ? ring : 0.0);
          }
          K[3 * i + k] = acc;
        }
      }
#pragma unroll
      for (int m = 0; m < 9; ++m) out[9 * (a + 3 * b) + m] = K[m];
    }
}

template <int MODE, class T>
__global__ void __launch_bounds__(256) wall_lhs_add(int nwk, const int *__restrict__ wk, const int *__restrict__ wptr,
                                                    const int *__restrict__ wsrc, const double *__restrict__ wrec,
                                                    T *__restrict__ out, size_t nnz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nwk) return;
  const size_t k = (size_t)wk[t];
  double s[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int p = wptr[t]; p < wptr[t + 1]; ++p) {
    const double *r = wrec + (size_t)wsrc[p] * 9;
#pragma unroll
    for (int m = 0; m < 9; ++m) s[m] += r[m];
  }
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    T *o = MODE == 0 ? out + 9 * k + m : out + (size_t)(m + m / 3) * nnz + k;
    *o = (T)((double)*o + s[m]);
  }
}

__global__ void __launch_bounds__(256) wall_rows_add(int nwf, const int *__restrict__ wface,
                                                     const double *__restrict__ wres, double *__restrict__ brec) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwf) return;
  double *b = brec + (size_t)wface[w] * 12;
  const double *r = wres + (size_t)w * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) b[4 * a + i] = b[4 * a + i] + r[3 * a + i];
}

// itrPredict's wall displacement (itrPC.f90:127-132)
__global__ void wall_predict(int nwn, int nshg, const int *__restrict__ wn, const double *__restrict__ uold,
                             const double *__restrict__ yold, const double *__restrict__ acold,
                             double *__restrict__ u, double Delt, double c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nwn) return;
  const size_t i = (size_t)wn[t];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    u[i + nshg * k] = uold[i + nshg * k] + Delt * yold[i + nshg * k] + c2 * acold[i + nshg * k];
}

// itrCorrect's (itrPC.f90:183-186): u += Delt^2 betai solinc, solinc = Res4^T
__global__ void wall_correct(int nwn, int nshg, const int *__restrict__ wn, double *__restrict__ u,
                             const double *__restrict__ res4, double f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nwn) return;
  const size_t i = (size_t)wn[t];
#pragma unroll
  for (int k = 0; k < 3; ++k) u[i + nshg * k] = u[i + nshg * k] + f * res4[4 * i + k];
}

// itrYAlpha's (itrPC.f90:333-336)
__global__ void wall_alpha(int nwn, int nshg, const int *__restrict__ wn, const double *__restrict__ uold,
                           const double *__restrict__ u, double *__restrict__ ua, double alfi) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nwn) return;
  const size_t i = (size_t)wn[t];
#pragma unroll
  for (int k = 0; k < 3; ++k) ua[i + nshg * k] = uold[i + nshg * k] + alfi * (u[i + nshg * k] - uold[i + nshg * k]);
}

int find_col(const std::vector<int> &rp, const std::vector<int> &cols, int i, int j) {
  const auto b = cols.begin() + rp[i], e = cols.begin() + rp[i + 1];
  const auto it = std::lower_bound(b, e, j);
  return (it != e && *it == j) ? (int)(it - cols.begin()) : -1;
}

}  // namespace

namespace crimson {
int form_wall_pass(crimson_form *f, int lhs, int mode, double *lhsK, double *V, float *V32, size_t nnz) {
  if (!f->wall || f->nwf == 0) return CRIMSON_OK;
  hipStream_t s = f->stream;
  const unsigned g = grid_for(f->nwf, 256);
  if (lhs)
    wall_elem<true><<<g, 256, 0, s>>>(f->nwf, f->d_wface.p, f->nelb, f->d_ienb.p, f->d_node.p, f->d_u, f->d_uref.p,
                                      f->nshg, f->d_wtag.p, f->nqb, f->wc, f->d_wres.p, f->d_wrec.p);
  else
    wall_elem<false><<<g, 256, 0, s>>>(f->nwf, f->d_wface.p, f->nelb, f->d_ienb.p, f->d_node.p, f->d_u, f->d_uref.p,
                                       f->nshg, f->d_wtag.p, f->nqb, f->wc, f->d_wres.p, f->d_wrec.p);
  if (lhs && f->nwk > 0) {
    const unsigned gk = grid_for(f->nwk, 256);
    if (mode == 0)
      wall_lhs_add<0, double><<<gk, 256, 0, s>>>(f->nwk, f->d_wk.p, f->d_wptr.p, f->d_wsrc.p, f->d_wrec.p, lhsK, nnz);
    else if (V32)
      wall_lhs_add<1, float><<<gk, 256, 0, s>>>(f->nwk, f->d_wk.p, f->d_wptr.p, f->d_wsrc.p, f->d_wrec.p, V32, nnz);
    else
      wall_lhs_add<1, double><<<gk, 256, 0, s>>>(f->nwk, f->d_wk.p, f->d_wptr.p, f->d_wsrc.p, f->d_wrec.p, V, nnz);
  }
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}
bool form_has_wall(const crimson_form *f) { return f->wall; }
int form_wall_step_ready(crimson_form *f) {
  if (!f->wall) return CRIMSON_OK;
  if (!f->d_wuold || !f->d_wu || !f->d_u) {
    set_error("a captured timestep with a deformable wall needs crimson_form_set_wall_state_dev and "
              "crimson_form_set_displacement_dev");
    return CRIMSON_ESTATE;
  }
  return CRIMSON_OK;
}
int form_wall_enqueue_predict(crimson_form *f, const double *d_yold, const double *d_acold) {
  if (!f->wall) return CRIMSON_OK;
  return crimson_wall_predict_dev(f, f->d_wuold, d_yold, d_acold, f->d_wu);
}
int form_wall_enqueue_alpha(crimson_form *f) {
  if (!f->wall) return CRIMSON_OK;
  return crimson_wall_alpha_dev(f, f->d_wuold, f->d_wu, const_cast<double *>(f->d_u));
}
int form_wall_enqueue_correct(crimson_form *f, const double *d_Res4) {
  if (!f->wall) return CRIMSON_OK;
  return crimson_wall_correct_dev(f, f->d_wu, d_Res4);
}
int form_wall_enqueue_update(crimson_form *f) {
  if (!f->wall) return CRIMSON_OK;
  CR_HIP(hipMemcpyAsync(f->d_wuold, f->d_wu, (size_t)3 * f->nshg * sizeof(double), hipMemcpyDeviceToDevice,
                        f->stream));
  return CRIMSON_OK;
}
int form_wall_rows(crimson_form *f) {
  if (!f->wall || f->nwf == 0) return CRIMSON_OK;
  wall_rows_add<<<grid_for(f->nwf, 256), 256, 0, f->stream>>>(f->nwf, f->d_wface.p, f->d_wres.p, f->d_brec.p);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}
}  // namespace crimson

extern "C" int crimson_form_set_wall(crimson_form *f, const crimson_wall_params *w, const int *nodetag,
                                     const double *uref) {
  if (!f || !w) return crimson::set_error("crimson_form_set_wall: null argument"), CRIMSON_EARG;
  if (f->prm.ideformwall != 1) {
    crimson::set_error("crimson_form_set_wall: the handle was set up with ideformwall = %d (1 needed)",
                       f->prm.ideformwall);
    return CRIMSON_ESTATE;
  }
  if (f->nelb == 0) {
    crimson::set_error("crimson_form_set_wall: no boundary elements (crimson_form_set_boundary first)");
    return CRIMSON_ESTATE;
  }
  if (!(w->thicknessvw > 0.0) || !(std::fabs(w->rnuvw) < 1.0) || w->iwallmassfactor < 0 ||
      w->iwallstiffactor < 0) {
    crimson::set_error("crimson_form_set_wall: thickness %g, Poisson ratio %g, mass/stiffness factors %d/%d "
                       "out of range",
                       w->thicknessvw, w->rnuvw, w->iwallmassfactor, w->iwallstiffactor);
    return CRIMSON_EARG;
  }
  const bool ring = w->iringsupp > 0 || w->iringdamp > 0;
  if (ring && !nodetag) {
    crimson::set_error("crimson_form_set_wall: the ring terms need the node tags");
    return CRIMSON_EARG;
  }
  hipStream_t s = f->stream;
  // the boundary elements (device copies) back on the host
  std::vector<int> ienb((size_t)4 * f->nelb), code((size_t)2 * f->nelb);
  std::vector<double> node;
  CR_HIP(hipMemcpyAsync(ienb.data(), f->d_ienb.p, ienb.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  CR_HIP(hipMemcpyAsync(code.data(), f->d_bcode.p, code.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  CR_HIP(hipStreamSynchronize(s));
  std::vector<int> wface;
  for (int e = 0; e < f->nelb; ++e)
    if (code[e] & 16) wface.push_back(e);
  const int nwf = (int)wface.size();
  // the wall nodes in the order gendat.f90:172-225 finds them
  std::vector<int> seen(f->nshg, 0), wnode;
  for (int e : wface)
    for (int a = 0; a < 3; ++a) {
      const int g = ienb[(size_t)a * f->nelb + e];
      if (!seen[g]) {
        seen[g] = 1;
        wnode.push_back(g);
      }
    }
  std::vector<int> wtag;
  if (ring) {
    wtag.resize((size_t)3 * nwf);
    for (int w2 = 0; w2 < nwf; ++w2) {
      int nr = 0;
      for (int a = 0; a < 3; ++a) {
        const int g = ienb[(size_t)a * f->nelb + wface[w2]];
        wtag[3 * (size_t)w2 + a] = nodetag[g] & 1;
        nr += nodetag[g] & 1;
      }
      if (nr == 3) {
        crimson::set_error("crimson_form_set_wall: wall face %d has three ring nodes (the tethering direction of "
                           "e3bvar.f90:909-925 is then undefined)",
                           wface[w2] + 1);
        return CRIMSON_EUNSUPPORTED;
      }
    }
  }
  // the tangent's gather: each face pair's block of the CSR graph
  std::vector<int> wk, wptr, wsrc;
  {
    const std::vector<int> &rp = f->h_wrp, &cols = f->h_wcols;
    std::vector<std::pair<int, int>> c;  // (block, source = face*9 + a + 3b)
    c.reserve((size_t)nwf * 9);
    for (int w2 = 0; w2 < nwf; ++w2)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) {
          const int ga = ienb[(size_t)a * f->nelb + wface[w2]], gb = ienb[(size_t)b * f->nelb + wface[w2]];
          const int k = find_col(rp, cols, ga, gb);
          if (k < 0) {
            crimson::set_error("crimson_form_set_wall: face pair (%d,%d) missing from the CSR graph", ga + 1, gb + 1);
            return CRIMSON_EGRAPH;
          }
          c.push_back({k, w2 * 9 + a + 3 * b});
        }
    std::stable_sort(c.begin(), c.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
      return x.first < y.first || (x.first == y.first && x.second < y.second);
    });
    for (size_t p = 0; p < c.size(); ++p) {
      if (p == 0 || c[p].first != c[p - 1].first) {
        wk.push_back(c[p].first);
        wptr.push_back((int)p);
      }
      wsrc.push_back(c[p].second);
    }
    wptr.push_back((int)c.size());
  }
  f->wall = true;
  f->wall_gen++;
  const crimson_form_params &p = f->prm;
  const double betai = 1.0;
  WallConst c{};
  c.h = w->thicknessvw;
  c.rhovw = w->rhovw;
  const double tc = 1.0 / (1.0 - w->rnuvw * w->rnuvw);
  c.d7 = w->evw * tc * 1.0;
  c.d8 = w->evw * tc * w->rnuvw;
  c.d9 = w->evw * tc * 0.5 * (1.0 - w->rnuvw);
  c.d10 = w->evw * tc * 0.5 * (1.0 - w->rnuvw) * w->rshearconstantvw;
  c.kstiff = w->iwallstiffactor * betai * p.Delt * p.Delt * p.alfi;
  c.mass = p.almi * (1.0 - p.flmpl) * w->rhovw * w->thicknessvw;
  c.supp = w->tissSuppStiffCoeff;
  c.supp_lhs = p.Delt * p.alfi * betai * p.Delt * w->tissSuppStiffCoeff;
  c.damp = w->tissSuppDampCoeff;
  c.damp_lhs = p.Delt * p.alfi * p.gami * w->tissSuppDampCoeff;
  c.ring_k = w->tissSuppRingStiffCoeff;
  c.ring_c = w->tissSuppRingDampCoeff;
  c.ring_klhs = betai * p.Delt * p.Delt * p.alfi;
  c.ring_clhs = p.Delt * p.alfi * p.gami;
  c.massf = w->iwallmassfactor;
  c.isupp = w->iwallsupp == 1;
  c.idamp = w->iwalldamp == 1;
  c.iring_k = w->iringsupp > 0;
  c.iring_c = w->iringdamp > 0;
  f->wc = c;
  f->nwf = nwf;
  f->nwn = (int)wnode.size();
  f->nwk = (int)wk.size();
  int rc;
  if ((rc = f->d_wface.upload(wface.data(), wface.size(), s)) ||
      (rc = f->d_wnode.upload(wnode.data(), wnode.size(), s)) ||
      (rc = f->d_wres.alloc((size_t)std::max(nwf, 1) * 9)) ||
      (rc = f->d_wrec.alloc((size_t)std::max(nwf, 1) * 81)) || (rc = f->d_wk.upload(wk.data(), wk.size(), s)) ||
      (rc = f->d_wptr.upload(wptr.data(), wptr.size(), s)) || (rc = f->d_wsrc.upload(wsrc.data(), wsrc.size(), s)))
    return f->wall = false, rc;
  if (ring) {
    if ((rc = f->d_wtag.upload(wtag.data(), wtag.size(), s))) return f->wall = false, rc;
  } else {
    f->d_wtag.release();
  }
  if (uref) {
    if ((rc = f->d_uref.upload(uref, (size_t)3 * f->nshg, s))) return f->wall = false, rc;
  } else {
    f->d_uref.release();
  }
  CR_HIP(hipStreamSynchronize(s));
  return CRIMSON_OK;
}

extern "C" int crimson_form_set_wall_state_dev(crimson_form *f, double *d_uold, double *d_u) {
  if (!f || !d_uold || !d_u) return crimson::set_error("crimson_form_set_wall_state_dev: null argument"), CRIMSON_EARG;
  f->d_wuold = d_uold;
  f->d_wu = d_u;
  f->wall_gen++;
  return CRIMSON_OK;
}

extern "C" int crimson_form_set_displacement_dev(crimson_form *f, const double *d_u) {
  if (!f || !d_u) return crimson::set_error("crimson_form_set_displacement_dev: null argument"), CRIMSON_EARG;
  f->d_u = d_u;
  return CRIMSON_OK;
}

namespace {
int wall_ready(crimson_form *f, const char *who) {
  if (!f) return crimson::set_error("%s: null handle", who), CRIMSON_EARG;
  if (!f->wall) {
    crimson::set_error("%s: no deformable wall (crimson_form_set_wall)", who);
    return CRIMSON_ESTATE;
  }
  return CRIMSON_OK;
}
}  // namespace

extern "C" int crimson_wall_predict_dev(crimson_form *f, const double *d_uold, const double *d_yold,
                                        const double *d_acold, double *d_u) {
  CR_TRY(wall_ready(f, "crimson_wall_predict_dev"));
  if (!d_uold || !d_yold || !d_acold || !d_u) return crimson::set_error("null array"), CRIMSON_EARG;
  if (f->nwn == 0) return CRIMSON_OK;
  const crimson_form_params &p = f->prm;
  const double betai = 1.0;
  const double c2 = 0.5 * ((p.gami - 2.0 * betai) / p.gami) * p.Delt * p.Delt;
  wall_predict<<<crimson::grid_for(f->nwn, 256), 256, 0, f->stream>>>(f->nwn, f->nshg, f->d_wnode.p, d_uold, d_yold,
                                                                      d_acold, d_u, p.Delt, c2);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_wall_correct_dev(crimson_form *f, double *d_u, const double *d_Res4) {
  CR_TRY(wall_ready(f, "crimson_wall_correct_dev"));
  if (!d_u || !d_Res4) return crimson::set_error("null array"), CRIMSON_EARG;
  if (f->nwn == 0) return CRIMSON_OK;
  const double betai = 1.0;
  wall_correct<<<crimson::grid_for(f->nwn, 256), 256, 0, f->stream>>>(f->nwn, f->nshg, f->d_wnode.p, d_u, d_Res4,
                                                                      f->prm.Delt * f->prm.Delt * betai);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}

extern "C" int crimson_wall_alpha_dev(crimson_form *f, const double *d_uold, const double *d_u, double *d_uAlpha) {
  CR_TRY(wall_ready(f, "crimson_wall_alpha_dev"));
  if (!d_uold || !d_u || !d_uAlpha) return crimson::set_error("null array"), CRIMSON_EARG;
  if (f->nwn == 0) return CRIMSON_OK;
  wall_alpha<<<crimson::grid_for(f->nwn, 256), 256, 0, f->stream>>>(f->nwn, f->nshg, f->d_wnode.p, d_uold, d_u,
                                                                    d_uAlpha, f->prm.alfi);
  CR_HIP(hipGetLastError());
  return CRIMSON_OK;
}
